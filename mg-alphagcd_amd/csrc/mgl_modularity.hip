// Modularity of a partition of a device-resident graph (the reference's
// compute_modularity, src/louvain/louvain.cu:105-238, for any assignment):
//   k_v   = sum of row v's weights (self-loops included, louvain.cu:30-74)
//   a_c   = sum of k_v over the members of c
//   2m    = sum of k_v                                 (host_graph.cpp:112-114)
//   Q     = sum_in / 2m - sum_c a_c^2 / (2m)^2          (louvain.cu:146-165, 187-206, 234)
// where sum_in adds the weights of the entries whose endpoints share a
// community. Unit weights (no weight array) are summed exactly in integers
// (int64 totals, 128-bit sum of squares, one rounding per term of Q); float
// weights in FP64.
//
// The inputs are the caller's device arrays and are checked on the device:
// a community id or column outside [0, n) or a row outside [0, nnz) is
// skipped and reported (MGL_INVALID_GRAPH), so no access leaves the arrays.
#include "mgl_kernels.h"
#include "../../include/mgl.h"

#include <cmath>
#include <cstdio>
#include <mutex>
#include <string>
#include <type_traits>

// Per-call working memory: a header of sums and flags, then the n community
// totals (int64 for unit weights, double otherwise).
struct ModHeader {
  int64_t sin_i;           // sum_in, unit weights
  int64_t mass_i;          // 2m, unit weights
  uint64_t sa2_lo, sa2_hi; // sum of a_c^2, unit weights (128-bit)
  double sin_d, mass_d, sa2_d;
  int32_t err;             // 1: community id out of range, 2: column out of range,
                           // 4: row outside the entry array
  int32_t pad;
};
static_assert(sizeof(ModHeader) <= 256, "header fits its 256-byte slot");
static const size_t MOD_HEADER_BYTES = 256;

static size_t mod_scratch_bytes(int64_t n) { return MOD_HEADER_BYTES + 8 * (size_t)(n > 0 ? n : 1); }

#define MOD_LANES 16  // lanes per vertex

__device__ __forceinline__ void add_total(int64_t* p, int64_t x) {
  atomicAdd((unsigned long long*)p, (unsigned long long)x);
}
__device__ __forceinline__ void add_total(double* p, double x) { atomicAdd(p, x); }

template <bool UNIT>
__global__ __launch_bounds__(256) void k_mod_rows(int64_t n, int64_t nnz,
                                                  const int64_t* __restrict__ off,
                                                  const int32_t* __restrict__ col,
                                                  const float* __restrict__ w,
                                                  const int32_t* __restrict__ c,
                                                  unsigned char* __restrict__ scr) {
  typedef typename std::conditional<UNIT, int64_t, double>::type T;
  ModHeader* H = (ModHeader*)scr;
  T* a = (T*)(scr + MOD_HEADER_BYTES);
  const int lane = threadIdx.x & (MOD_LANES - 1);
  const int64_t grp0 = ((int64_t)blockIdx.x * 256 + threadIdx.x) / MOD_LANES;
  const int64_t ngrp = ((int64_t)gridDim.x * 256) / MOD_LANES;
  T sin = 0, mass = 0;
  int err = 0;
  for (int64_t v = grp0; v < n; v += ngrp) {
    const int32_t cv = c[v];
    const bool cok = cv >= 0 && cv < n;
    const int64_t e0 = off[v], e1 = off[v + 1];
    T k = 0, s = 0;
    if (e0 < 0 || e1 < e0 || e1 > nnz) {
      err |= 4;
    } else {
      for (int64_t e = e0 + lane; e < e1; e += MOD_LANES) {
        const int32_t u = col[e];
        const T x = UNIT ? (T)1 : (T)w[e];
        k += x;
        if (u < 0 || u >= n) { err |= 2; continue; }
        const int32_t cu = c[u];
        if (cok && cu == cv) s += x;
      }
    }
    if (!cok) err |= 1;
    k = group_sum<MOD_LANES>(k);
    s = group_sum<MOD_LANES>(s);
    if (lane == 0 && cok) add_total(&a[cv], k);
    if (lane == 0) {
      sin += s;
      mass += k;
    }
  }
  if (err) atomicOr(&H->err, err);  // malformed input only
  sin = group_sum<64>(sin);
  mass = group_sum<64>(mass);
  if ((threadIdx.x & 63) == 0) {
    if constexpr (UNIT) {
      if (sin) add_total(&H->sin_i, sin);
      if (mass) add_total(&H->mass_i, mass);
    } else {
      if (sin != 0) add_total(&H->sin_d, sin);
      if (mass != 0) add_total(&H->mass_d, mass);
    }
  }
}

__device__ __forceinline__ u128 shfl_xor_u128(u128 x, int m) {
  const uint64_t lo = (uint64_t)x, hi = (uint64_t)(x >> 64);
  const uint64_t l2 = (uint64_t)shfl_xor_t((int64_t)lo, m, 64);
  const uint64_t h2 = (uint64_t)shfl_xor_t((int64_t)hi, m, 64);
  return ((u128)h2 << 64) | l2;
}

template <bool UNIT>
__global__ __launch_bounds__(256) void k_mod_sa2(int64_t n, unsigned char* __restrict__ scr) {
  ModHeader* H = (ModHeader*)scr;
  if constexpr (UNIT) {
    const int64_t* a = (const int64_t*)(scr + MOD_HEADER_BYTES);
    u128 s = 0;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * 256) {
      const int64_t x = a[i];
      const u128 ax = (u128)(x < 0 ? -x : x);
      s += ax * ax;
    }
#pragma unroll
    for (int m = 32; m > 0; m >>= 1) s += shfl_xor_u128(s, m);
    if ((threadIdx.x & 63) == 0 && s) atomic_add_u128(&H->sa2_lo, &H->sa2_hi, s);
  } else {
    const double* a = (const double*)(scr + MOD_HEADER_BYTES);
    double s = 0;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * 256)
      s += a[i] * a[i];
    s = group_sum<64>(s);
    if ((threadIdx.x & 63) == 0 && s != 0) atomicAdd(&H->sa2_d, s);
  }
}

static unsigned mod_grid(int64_t n, int per_block) {
  int64_t g = (n + per_block - 1) / per_block;
  if (g < 1) g = 1;
  if (g > 8192) g = 8192;
  return (unsigned)g;
}

// Runs the computation with working memory scr (mod_scratch_bytes(n) bytes)
// on stream st and returns Q. err receives the input error flags.
static hipError_t mod_run(int64_t n, int64_t nnz, const int64_t* off, const int32_t* col,
                          const float* w, const int32_t* c, unsigned char* scr, hipStream_t st,
                          double* q, int* err) {
  // everything that accumulates starts at zero: the header and the totals
  launch_fill_bytes0(scr, mod_scratch_bytes(n), st);
  if (n > 0) {
    const unsigned gr = mod_grid(n, 256 / MOD_LANES);
    if (w) k_mod_rows<false><<<gr, 256, 0, st>>>(n, nnz, off, col, w, c, scr);
    else k_mod_rows<true><<<gr, 256, 0, st>>>(n, nnz, off, col, w, c, scr);
    if (w) k_mod_sa2<false><<<mod_grid(n, 256), 256, 0, st>>>(n, scr);
    else k_mod_sa2<true><<<mod_grid(n, 256), 256, 0, st>>>(n, scr);
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  ModHeader h;
  e = hipMemcpyAsync(&h, scr, sizeof(h), hipMemcpyDeviceToHost, st);
  if (e != hipSuccess) return e;
  e = hipStreamSynchronize(st);
  if (e != hipSuccess) return e;
  *err = h.err;
  if (!w) {
    const double M = (double)h.mass_i;
    *q = (h.mass_i == 0) ? 0.0 : (double)h.sin_i / M - u128_to_double(h.sa2_hi, h.sa2_lo) / (M * M);
  } else {
    const double M = h.mass_d;
    *q = (M == 0) ? 0.0 : h.sin_d / M - h.sa2_d / (M * M);
  }
  return hipSuccess;
}

static mgl_status mod_args(mgl_ctx* ctx, int64_t n, int64_t nnz, const int64_t* off,
                           const int32_t* col, const int32_t* c, const double* q,
                           std::string& msg) {
  (void)ctx;
  if (!q) { msg = "null output"; return MGL_INVALID_ARG; }
  if (n < 0 || nnz < 0) { msg = "negative size"; return MGL_INVALID_ARG; }
  if (n > 0x7fffffffLL) { msg = "more than 2^31-1 vertices"; return MGL_UNSUPPORTED; }
  if (n > 0 && (!off || !c || (nnz > 0 && !col))) { msg = "null device array"; return MGL_INVALID_ARG; }
  return MGL_OK;
}

static mgl_status mod_result(int err, std::string& msg) {
  if (!err) return MGL_OK;
  msg = "invalid input:";
  if (err & 1) msg += " community id outside [0, n)";
  if (err & 2) msg += " column outside [0, n)";
  if (err & 4) msg += " row outside [0, nnz)";
  return MGL_INVALID_GRAPH;
}

// defined in mgl_engine.hip: sets the context's last-error string
void mgl_set_error(mgl_ctx* ctx, const std::string& msg);
int mgl_ctx_device(const mgl_ctx* ctx);
hipStream_t mgl_ctx_stream(const mgl_ctx* ctx);

extern "C" mgl_status mgl_modularity_dev(mgl_ctx* ctx, int64_t n, int64_t nnz,
                                         const int64_t* d_offsets, const int32_t* d_cols,
                                         const float* d_w, const int32_t* d_assign, double* q) {
  if (!ctx) return MGL_INVALID_ARG;
  std::string msg;
  mgl_status s = mod_args(ctx, n, nnz, d_offsets, d_cols, d_assign, q, msg);
  if (s) { mgl_set_error(ctx, msg); return s; }
  if (hipSetDevice(mgl_ctx_device(ctx)) != hipSuccess) {
    (void)hipGetLastError();
    mgl_set_error(ctx, "hipSetDevice failed");
    return MGL_DEVICE_ERROR;
  }
  void* scr = nullptr;
  hipError_t e = hipMalloc(&scr, mod_scratch_bytes(n));
  if (e != hipSuccess) {
    (void)hipGetLastError();
    mgl_set_error(ctx, std::string("hipMalloc: ") + hipGetErrorString(e));
    return e == hipErrorOutOfMemory ? MGL_OOM : MGL_DEVICE_ERROR;
  }
  int err = 0;
  e = mod_run(n, nnz, d_offsets, d_cols, d_w, d_assign, (unsigned char*)scr, mgl_ctx_stream(ctx),
              q, &err);
  (void)hipFree(scr);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    mgl_set_error(ctx, std::string("modularity: ") + hipGetErrorString(e));
    return MGL_DEVICE_ERROR;
  }
  s = mod_result(err, msg);
  if (s) mgl_set_error(ctx, msg);
  return s;
}

// ---------------------------------------------------------------------------
// Allocation-free variant: the same kernels with their working memory in a
// file-scope device array. Each device the library runs on holds its own copy
// for the life of the process; its address on a device comes from
// hipGetSymbolAddress, taken once per device with that device current (at
// mgl_create). Host code never passes the array's name itself to a kernel.
static __device__ __align__(256) unsigned char
    scratch[MOD_HEADER_BYTES + 8 * (size_t)MGL_NOALLOC_MAX_VERTICES];
static_assert(MGL_NOALLOC_MAX_VERTICES <= 0x7fffffff, "vertex limit fits int32 ids");

#define MOD_MAX_DEVICES 64
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[MOD_MAX_DEVICES];  // device address of scratch per device
static int g_scratch_state[MOD_MAX_DEVICES];  // 0 not looked up, 1 ok, -1 failed

// Looks up the array's address on `device`, which must be the current device.
int mgl_noalloc_setup(int device) {
  if (device < 0 || device >= MOD_MAX_DEVICES) return -1;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_state[device] != 0) return g_scratch_state[device] > 0 ? 0 : -1;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress((void**)&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    (void)hipGetLastError();
    g_scratch_state[device] = -1;
    return -1;
  }
  g_scratch_ptr[device] = p;
  g_scratch_state[device] = 1;
  return 0;
}

extern "C" mgl_status mgl_modularity_dev_noalloc(mgl_ctx* ctx, int64_t n, int64_t nnz,
                                                 const int64_t* d_offsets,
                                                 const int32_t* d_cols, const float* d_w,
                                                 const int32_t* d_assign, double* q) {
  if (!ctx) return MGL_INVALID_ARG;
  std::string msg;
  mgl_status s = mod_args(ctx, n, nnz, d_offsets, d_cols, d_assign, q, msg);
  if (s) { mgl_set_error(ctx, msg); return s; }
  if (n > MGL_NOALLOC_MAX_VERTICES) {
    char b[160];
    snprintf(b, sizeof(b), "%lld vertices exceed the %d the allocation-free entry point holds",
             (long long)n, (int)MGL_NOALLOC_MAX_VERTICES);
    mgl_set_error(ctx, b);
    return MGL_INVALID_ARG;
  }
  const int dev = mgl_ctx_device(ctx);
  if (hipSetDevice(dev) != hipSuccess) {
    (void)hipGetLastError();
    mgl_set_error(ctx, "hipSetDevice failed");
    return MGL_DEVICE_ERROR;
  }
  if (mgl_noalloc_setup(dev) != 0) {
    mgl_set_error(ctx, "the allocation-free working memory has no address on this device");
    return MGL_DEVICE_ERROR;
  }
  unsigned char* scr;
  {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    scr = (unsigned char*)g_scratch_ptr[dev];
  }
  int err = 0;
  const hipError_t e =
      mod_run(n, nnz, d_offsets, d_cols, d_w, d_assign, scr, mgl_ctx_stream(ctx), q, &err);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    mgl_set_error(ctx, std::string("modularity: ") + hipGetErrorString(e));
    return MGL_DEVICE_ERROR;
  }
  s = mod_result(err, msg);
  if (s) mgl_set_error(ctx, msg);
  return s;
}
