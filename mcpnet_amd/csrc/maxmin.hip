// Max-min ("tropical") products C[i][j] = max_k min(A[i][k], B[j][k]), the guarded ratio and the
// MCP2/3/4 pipelines built from them.
//
// The product is a semiring contraction with no rounding, so it runs on the VALU (MFMA has no
// max/min form).  FP32 values are mapped to int32 keys whose signed order equals the float
// order (negative floats get their magnitude bits flipped); the inner loop is then one
// v_min_i32 per triple plus one v_max3_i32 per two triples, and the result is exact for every
// tiling and reduction order.  FP64 uses v_min_f64 / v_max_f64.
//
// Tiling: a 256-thread workgroup owns a 128 x 128 output tile; each thread keeps an 8 x 8 block
// of accumulators (rows ty+16m, cols tx+16n).
// * FP32: the operands are converted once to padded int32 key matrices (key_prep_kernel; the K
//   mask of the masked kernel is folded in there).  K advances in chunks of 16 keys per row that
//   arrive in LDS by direct buffer loads, two chunks in flight, into XOR-swizzled rows whose
//   operand reads are bank-conflict free.  An MCP pipeline keeps M1 in key form on the device.
// * FP64: the operands are copied once to padded value matrices (val_prep_kernel, -inf padding,
//   K mask folded in) and K chunks of 16 doubles per row arrive in LDS by the same direct loads
//   (maxmin_f64_kernel).  The earlier register-staged kernel (maxmin_tile_kernel: chunks of 128
//   bytes staged through registers into LDS rows padded to 144 bytes) stays as a cross-check.
#include <cfloat>
#include <climits>
#include <cmath>
#include <cstdio>
#include <limits>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "internal.h"

namespace {

using mcpg::ceil_div;

// ---------------------------------------------------------------- element traits
struct TrF32 {
  using T = float;
  using E = int;                                   // LDS / register element (order key)
  using V = int4;                                  // 16-byte vector of E
  static constexpr int kVec = 4;
  __device__ static __forceinline__ E key(float x) {
    int b = __float_as_int(x);
    return b ^ ((b >> 31) & 0x7fffffff);
  }
  __device__ static __forceinline__ float val(E k) {
    int b = k ^ ((k >> 31) & 0x7fffffff);
    return __int_as_float(b);
  }
  __device__ static __forceinline__ E pad() { return INT_MIN; }      // neutral: below every key
  __device__ static __forceinline__ E init() { return key(-INFINITY); }  // (float)lowest<double>() == -inf
  __device__ static __forceinline__ E mn(E a, E b) { return a < b ? a : b; }
  __device__ static __forceinline__ E mx(E a, E b) { return a > b ? a : b; }
  // max(a, b, c) as one v_max3_i32.  Written as a single VALU instruction because the compiler
  // otherwise reassociates max(acc, max(m0, m1)) chains into two-input maxima (1.75 instead of
  // 1.5 instructions per (i, k, j) triple).
  __device__ static __forceinline__ E max3(E a, E b, E c) {
    E r;
    asm("v_max3_i32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
    return r;
  }
  // acc = max(acc, min(a.i, b.i)) over the 4 lanes of a 16-byte chunk
  __device__ static __forceinline__ E step(E acc, const V& a, const V& b) {
    E m0 = mn(a.x, b.x), m1 = mn(a.y, b.y), m2 = mn(a.z, b.z), m3 = mn(a.w, b.w);
    acc = max3(acc, m0, m1);
    acc = max3(acc, m2, m3);
    return acc;
  }
  __device__ static __forceinline__ V vmin(const V& a, const V& b) {
    return V{mn(a.x, b.x), mn(a.y, b.y), mn(a.z, b.z), mn(a.w, b.w)};
  }
};

struct TrF64 {
  using T = double;
  using E = double;
  using V = double2;
  static constexpr int kVec = 2;
  __device__ static __forceinline__ E key(double x) { return x; }
  __device__ static __forceinline__ double val(E k) { return k; }
  __device__ static __forceinline__ E pad() { return -INFINITY; }
  __device__ static __forceinline__ E init() { return -DBL_MAX; }  // numeric_limits<double>::lowest()
  __device__ static __forceinline__ E mn(E a, E b) { return __builtin_fmin(a, b); }
  __device__ static __forceinline__ E mx(E a, E b) { return __builtin_fmax(a, b); }
  __device__ static __forceinline__ E step(E acc, const V& a, const V& b) {
    return mx(acc, mx(mn(a.x, b.x), mn(a.y, b.y)));
  }
  __device__ static __forceinline__ V vmin(const V& a, const V& b) { return V{mn(a.x, b.x), mn(a.y, b.y)}; }
};

constexpr int kBM = 128;          // tile rows
constexpr int kBN = 128;          // tile cols
constexpr int kThreads = 256;
constexpr int kRowBytes = 128;    // K chunk per row, bytes
constexpr int kLdsRowBytes = 144; // padded LDS row
constexpr int kChunks = kRowBytes / 16;

// ratio_kernel semantics (combine.hpp:214-220)
template <typename T>
__device__ __forceinline__ T ratio_of(T a, T m) {
  const T eps = sizeof(T) == 4 ? (T)FLT_EPSILON : (T)DBL_EPSILON;
  const T big = sizeof(T) == 4 ? (T)FLT_MAX : (T)DBL_MAX;
  if (fabs(a) < eps) return (T)0;
  if (fabs(m) < eps) return big;
  return (T)((double)a / (double)m);
}

struct MMArgs {
  const void* A; int64_t lda;
  const void* B; int64_t ldb;
  const void* kmask;       // nullable
  int64_t M, N, K;
  void* C; int64_t ldc;    // nullable (ratio-only output)
  const void* R; int64_t ldr;  // ratio numerator (nullable: no fused ratio)
  void* O; int64_t ldo;        // ratio output
  int tri;                 // compute only tiles bm <= bn and mirror them (result known symmetric)
  int shared;              // A and B are the same matrix (diagonal tiles stage one panel)
  int aligned;             // rows of A and B are 16-byte aligned and lda/ldb multiples of 16 bytes
  int64_t nbm, nbn, ntiles;
  const uint32_t* order;   // tri mode: tile coordinates in launch order, (bm << 16) | bn
};

template <class Tr>
__device__ __forceinline__ void load_chunk(const typename Tr::T* __restrict__ P, int64_t ld, int64_t rows,
                                           int64_t K, int64_t row0, int64_t k0, int aligned,
                                           const typename Tr::T* __restrict__ kmask, int tid,
                                           typename Tr::V (&st)[4]) {
  using E = typename Tr::E;
  using T = typename Tr::T;
  constexpr int kv = Tr::kVec;
  const int c = tid & 7;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int64_t r = row0 + (tid >> 3) + 32 * q;
    const int64_t k = k0 + (int64_t)c * kv;
    E e[kv];
    if (aligned && r < rows && k + kv <= K) {
      typename Tr::V raw = *reinterpret_cast<const typename Tr::V*>(P + r * ld + k);
      const T* rv = reinterpret_cast<const T*>(&raw);
#pragma unroll
      for (int i = 0; i < kv; ++i) e[i] = Tr::key(rv[i]);
    } else {
#pragma unroll
      for (int i = 0; i < kv; ++i) e[i] = (r < rows && k + i < K) ? Tr::key(P[r * ld + k + i]) : Tr::pad();
    }
    if (kmask) {
#pragma unroll
      for (int i = 0; i < kv; ++i)
        if (k + i < K) e[i] = Tr::mn(e[i], Tr::key(kmask[k + i]));
    }
    st[q] = *reinterpret_cast<typename Tr::V*>(e);
  }
}

template <class Tr>
__global__ __launch_bounds__(kThreads) void maxmin_tile_kernel(MMArgs a) {
  using E = typename Tr::E;
  using V = typename Tr::V;
  using T = typename Tr::T;
  constexpr int kv = Tr::kVec;
  constexpr int kLdsRow = kLdsRowBytes / sizeof(E);

  __shared__ __align__(16) E As[kBM * kLdsRow];
  __shared__ __align__(16) E Bs[kBN * kLdsRow];

  const int64_t L = mcpg::xcd_remap(blockIdx.x, a.ntiles);
  int64_t bm, bn;
  if (a.tri) {
    const uint32_t code = a.order[blockIdx.x];
    bm = code >> 16;
    bn = code & 0xffffu;
  } else {
    const int64_t group = 8;
    const int64_t per_group = group * a.nbn;
    const int64_t g = L / per_group;
    const int64_t first = g * group;
    const int64_t gsize = (a.nbm - first) < group ? (a.nbm - first) : group;
    const int64_t in = L % per_group;
    bm = first + in % gsize;
    bn = in / gsize;
  }
  const int64_t row0 = bm * kBM, col0 = bn * kBN;
  const bool same = a.shared && bm == bn;

  const int tid = threadIdx.x;
  const int tx = tid & 15, ty = tid >> 4;
  const T* __restrict__ A = static_cast<const T*>(a.A);
  const T* __restrict__ B = static_cast<const T*>(a.B);
  const T* __restrict__ km = static_cast<const T*>(a.kmask);

  E acc[8][8];
#pragma unroll
  for (int m = 0; m < 8; ++m)
#pragma unroll
    for (int n = 0; n < 8; ++n) acc[m][n] = Tr::init();

  V sa[4], sb[4];
  constexpr int kChunkElems = kRowBytes / sizeof(E);
  const int64_t nk = ceil_div(a.K, kChunkElems);
  if (nk > 0) {
    load_chunk<Tr>(A, a.lda, a.M, a.K, row0, 0, a.aligned, km, tid, sa);
    if (!same) load_chunk<Tr>(B, a.ldb, a.N, a.K, col0, 0, a.aligned, nullptr, tid, sb);
  }
  for (int64_t kc = 0; kc < nk; ++kc) {
    // registers -> LDS
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int r = (tid >> 3) + 32 * q;
      *reinterpret_cast<V*>(&As[r * kLdsRow + (tid & 7) * kv]) = sa[q];
      if (!same) *reinterpret_cast<V*>(&Bs[r * kLdsRow + (tid & 7) * kv]) = sb[q];
    }
    __syncthreads();
    if (kc + 1 < nk) {  // prefetch the next chunk while this one is consumed
      load_chunk<Tr>(A, a.lda, a.M, a.K, row0, (kc + 1) * kChunkElems, a.aligned, km, tid, sa);
      if (!same) load_chunk<Tr>(B, a.ldb, a.N, a.K, col0, (kc + 1) * kChunkElems, a.aligned, nullptr, tid, sb);
    }
    const E* Bsrc = same ? As : Bs;
#pragma unroll 2
    for (int c = 0; c < kChunks; ++c) {
      V av[8], bv[8];
#pragma unroll
      for (int m = 0; m < 8; ++m) av[m] = *reinterpret_cast<const V*>(&As[(ty + 16 * m) * kLdsRow + c * kv]);
#pragma unroll
      for (int n = 0; n < 8; ++n) bv[n] = *reinterpret_cast<const V*>(&Bsrc[(tx + 16 * n) * kLdsRow + c * kv]);
#pragma unroll
      for (int m = 0; m < 8; ++m)
#pragma unroll
        for (int n = 0; n < 8; ++n) acc[m][n] = Tr::step(acc[m][n], av[m], bv[n]);
    }
    __syncthreads();
  }

  // epilogue: outputs, optional fused ratio, mirror for symmetric off-diagonal tiles
  T* __restrict__ C = static_cast<T*>(a.C);
  const T* __restrict__ R = static_cast<const T*>(a.R);
  T* __restrict__ O = static_cast<T*>(a.O);
  const bool mirror = a.tri && bm != bn;
#pragma unroll
  for (int m = 0; m < 8; ++m) {
    const int64_t i = row0 + ty + 16 * m;
    if (i >= a.M) continue;
#pragma unroll
    for (int n = 0; n < 8; ++n) {
      const int64_t j = col0 + tx + 16 * n;
      if (j >= a.N) continue;
      const T v = Tr::val(acc[m][n]);
      if (C) {
        C[i * a.ldc + j] = v;
        if (mirror) C[j * a.ldc + i] = v;
      }
      if (O) {
        O[i * a.ldo + j] = ratio_of<T>(R[i * a.ldr + j], v);
        if (mirror) O[j * a.ldo + i] = ratio_of<T>(R[j * a.ldr + i], v);
      }
    }
  }
}

// ---------------------------------------------------------------- FP32 key path
// The FP32 products run on int32 order keys prepared once per operand: rows padded to a multiple
// of the tile, row length ldk a multiple of the K chunk, padding INT_MIN (below every key, so it
// never wins a max of mins).  The K mask of the masked kernel is folded in at preparation.
__device__ __forceinline__ int f32_key(float x) {
  const int b = __float_as_int(x);
  return b ^ ((b >> 31) & 0x7fffffff);
}
__device__ __forceinline__ float key_f32(int k) { return __int_as_float(k ^ ((k >> 31) & 0x7fffffff)); }
// A key product widened to the FP64 kernel's value.  Its running maximum starts at
// lowest<double>() (simd_maxmin.hpp:32) where the key kernel's starts at -inf; with finite
// operands a result of -inf means that every K term was masked by a mask value of -inf or
// lowest<double>(), where the FP64 product gives lowest<double>().
__device__ __forceinline__ double key_f64(int k) {
  const float v = key_f32(k);
  return v == -INFINITY ? -DBL_MAX : (double)v;
}

// Keys of an FP64 operand, with flags[0] |= 1 unless every entry is a finite double that a float
// holds exactly and is not -0 (the one value on which the FP32 and FP64 kernels break ties
// differently).  A K-mask value above FLT_MAX acts as +inf and one of -inf or lowest<double>() as
// -inf (exact against finite operands: the first never wins a min, the second loses every max
// to an unmasked term); any other mask value must be held exactly too.
__global__ void key_prep_f64_kernel(const double* __restrict__ X, int64_t ldx, int64_t rows, int64_t K,
                                    const double* __restrict__ kmask, int* __restrict__ out, int64_t ldk,
                                    int64_t rows_pad, int* __restrict__ flags) {
  const int64_t total = rows_pad * ldk;
  bool bad = false;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = t / ldk, k = t % ldk;
    int v = INT_MIN;
    if (r < rows && k < K) {
      const double x = X[r * ldx + k];
      const float f = (float)x;
      bad |= !((double)f == x) || isinf(f) || (x == 0.0 && signbit(x));
      v = f32_key(f);
      if (kmask) {
        const double m = kmask[k];
        float fm = (float)m;
        if (m > (double)FLT_MAX) fm = INFINITY;
        else if (m == -INFINITY || m == -DBL_MAX) fm = -INFINITY;
        else bad |= !((double)fm == m) || isinf(fm) || (m == 0.0 && signbit(m));
        v = min(v, f32_key(fm));
      }
    }
    out[t] = v;
  }
  if (__ballot(bad) != 0 && (threadIdx.x & 63) == 0) atomicOr(flags, 1);
}

__global__ void key_prep_kernel(const float* __restrict__ X, int64_t ldx, int64_t rows, int64_t K,
                                const float* __restrict__ kmask, int* __restrict__ out, int64_t ldk, int64_t rows_pad) {
  const int64_t total = rows_pad * ldk;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = t / ldk, k = t % ldk;
    int v = INT_MIN;
    if (r < rows && k < K) {
      v = f32_key(X[r * ldx + k]);
      if (kmask) v = min(v, f32_key(kmask[k]));
    }
    out[t] = v;
  }
}

// INT_MIN into the padding of a key matrix whose [0, rows) x [0, cols) part a product writes
__global__ void key_pad_kernel(int* __restrict__ out, int64_t ldk, int64_t rows, int64_t cols, int64_t rows_pad) {
  const int64_t pad_cols = ldk - cols;
  const int64_t total = rows * pad_cols + (rows_pad - rows) * ldk;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
    if (t < rows * pad_cols) out[(t / pad_cols) * ldk + cols + t % pad_cols] = INT_MIN;
    else out[rows * ldk + (t - rows * pad_cols)] = INT_MIN;
  }
}

// DPI operand keys (include/mcp/filter/dpi.hpp:56-66): the diagonal is excluded (a triangle
// x-y-z needs y != x, z), and with a TF list the columns of non-TF genes too (for the TF
// variant's pairs with exactly one TF endpoint, dpi.hpp:90-101).  Excluded entries are INT_MIN.
template <typename T>
__global__ void dpi_key_kernel(const T* __restrict__ A, int64_t lda, int64_t N, const T* __restrict__ tfs,
                               int* __restrict__ out, int64_t ldk, int64_t rows_pad, int* __restrict__ flags) {
  const int64_t total = rows_pad * ldk;
  bool bad = false;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = t / ldk, k = t % ldk;
    int v = INT_MIN;
    if (r < N && k < N && r != k && !(tfs && !(tfs[k] > (T)0))) {
      const T x = A[r * lda + k];
      const float f = (float)x;
      if constexpr (sizeof(T) == 8)   // FP64: the key test of key_prep_f64_kernel on the entries used
        bad |= !((double)f == x) || isinf(f) || (x == 0.0 && signbit(x));
      v = f32_key(f);
    }
    out[t] = v;
  }
  if (flags && __ballot(bad) != 0 && (threadIdx.x & 63) == 0) atomicOr(flags, 1);
}

struct MKArgs {
  const int* A; int64_t lda;   // keys, rows padded to kBM, lda multiple of the K chunk
  const int* B; int64_t ldb;
  uint32_t a_bytes, b_bytes;   // buffer ranges
  int64_t M, N, Kp;            // Kp: padded K (multiple of the K chunk)
  float* C; int64_t ldc;       // float output (nullable)
  int* Ck; int64_t ldck;       // key output (nullable)
  const float* R; int64_t ldr; // ratio numerator (nullable: no fused ratio)
  float* O; int64_t ldo;       // ratio output
  // double forms of C, R and O (same leading dimensions; set instead of C / R / O): products of
  // FP64 operands that every float holds exactly, widened, with the ratio taken in double
  double* Cd;
  const double* Rd;
  double* Od;
  int tri, shared;
  int64_t nbm, nbn, ntiles;
  const uint32_t* order;       // tri mode: tile coordinates in launch order, (bm << 16) | bn
  // split tail: tiles nfull .. ntiles - 1 run as `splits` workgroups each, over K chunks
  // [part * kc_per, (part + 1) * kc_per), and fold their keys into tail (one 128 x 128 block of
  // keys per tile) with atomicMax; mm_finish_tail_kernel then writes their outputs
  int64_t nfull;
  int splits;
  int64_t kc_per;
  int* tail;
};

// tile -> (bm, bn): the packed tri order, or the 8-row grouped order of a full grid
__device__ __forceinline__ void mm_tile_coords(const MKArgs& a, int64_t t, int64_t& bm, int64_t& bn) {
  if (a.tri) {
    const uint32_t code = a.order[t];
    bm = code >> 16;
    bn = code & 0xffffu;
  } else {
    const int64_t group = 8;
    const int64_t per_group = group * a.nbn;
    const int64_t g = t / per_group;
    const int64_t first = g * group;
    const int64_t gsize = (a.nbm - first) < group ? (a.nbm - first) : group;
    const int64_t in = t % per_group;
    bm = first + in % gsize;
    bn = in / gsize;
  }
}

typedef __attribute__((address_space(3))) void* lds_ptr;
constexpr int kRsrcWord3 = 0x00020000;   // raw (untyped, unswizzled) gfx9-family buffer access
// LDS operand reads use a native vector type: with it the compiler leaves the ordering of the
// direct loads and the reads to the explicit waits and barriers of the pipeline (a struct type
// makes it wait for every pending load before each read, serialising the stages).
typedef int v4i __attribute__((ext_vector_type(4)));

__device__ __forceinline__ int step4(int acc, const v4i& a, const v4i& b) {
  const int m0 = min(a.x, b.x), m1 = min(a.y, b.y), m2 = min(a.z, b.z), m3 = min(a.w, b.w);
  acc = TrF32::max3(acc, m0, m1);
  return TrF32::max3(acc, m2, m3);
}

// K chunk of KC keys per stage.  LDS panel rows hold KC keys (KC / 4 chunks of 16 B) stored
// XOR-swizzled, chunk' = chunk ^ swz(row), so that the 16 distinct rows of a B-operand read
// (lane -> row lane & 15) and the two rows of an A-operand read fall in distinct bank quads of
// every 16-lane ds_read_b128 pass (passes: lanes {0-3, 12-15, 20-23, 24-27} and translates).
template <int KC>
struct KeyChunk {
  static constexpr int kRowBytes = KC * 4;
  static constexpr int kChunks = KC / 4;
  static constexpr int kRowsPerLoad = 1024 / kRowBytes;         // rows per 1 KB direct load
  static constexpr int kLoadsPerPanel = kBM / kRowsPerLoad;
  static constexpr int kPanelBytes = kBM * kRowBytes;
  static constexpr int kStageBytes = 2 * kPanelBytes;
  __device__ static __forceinline__ int swz(int row) {
    return KC == 32 ? (row >> 1) & 7 : (row >> 2) & 3;
  }
};

// this wave's direct loads of one K chunk: loads q0 .. q0 + nload - 1 of its panel
template <int KC>
__device__ __forceinline__ void key_dma_issue(__amdgpu_buffer_rsrc_t rsrc, const uint32_t (&lbyte)[2], uint32_t sbase,
                                              uint32_t rbytes, int nload, int q0, char* dst) {
#pragma unroll
  for (int j = 0; j < KeyChunk<KC>::kLoadsPerPanel / 2; ++j) {
    if (j < nload) {
      const int q = q0 + j;
      __builtin_amdgcn_raw_ptr_buffer_load_lds(rsrc, (lds_ptr)(dst + q * 1024), 16, lbyte[q & 1], sbase + (uint32_t)q * rbytes,
                                               0, 0);
    }
  }
}

// WIDE: the double outputs Cd / Od in place of C / O.  A separate instance: both epilogues in one
// took 182 VGPRs against 166, and the product one wave per SIMD fewer (2 against 3).
template <int KC, bool WIDE>
__global__ __launch_bounds__(kThreads) void maxmin_key_kernel(MKArgs a) {
  using KCh = KeyChunk<KC>;
  extern __shared__ __align__(16) char smem[];

  int64_t tile = blockIdx.x, part = -1;
  if (tile >= a.nfull) {
    part = (tile - a.nfull) % a.splits;
    tile = a.nfull + (tile - a.nfull) / a.splits;
  }
  int64_t bm, bn;
  mm_tile_coords(a, tile, bm, bn);
  const int64_t row0 = bm * kBM, col0 = bn * kBN;
  const bool same = a.shared && bm == bn;

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int tx = tid & 15, ty = tid >> 4;

  int acc[8][8];
#pragma unroll
  for (int m = 0; m < 8; ++m)
#pragma unroll
    for (int n = 0; n < 8; ++n) acc[m][n] = f32_key(-INFINITY);   // (float)lowest<double>() == -inf

  // Direct loads: waves 0-1 fill the A panel, waves 2-3 the B panel (on a diagonal tile of a
  // shared operand all four fill A), each a contiguous run of 1 KB loads.  A load covers
  // kRowsPerLoad rows; lane -> row lane / kChunks of the load, physical chunk lane % kChunks.
  const int swave = __builtin_amdgcn_readfirstlane(wave);
  const int panel = same ? 0 : (swave >> 1);
  const int nload = same ? KCh::kLoadsPerPanel / 4 : KCh::kLoadsPerPanel / 2;
  const int q0 = same ? swave * nload : (swave & 1) * nload;
  const int64_t prow0 = panel == 0 ? row0 : col0;
  const int64_t ld = panel == 0 ? a.lda : a.ldb;
  const __amdgpu_buffer_rsrc_t rsrc = __builtin_amdgcn_make_buffer_rsrc(
      (void*)(panel == 0 ? a.A : a.B), (short)0, panel == 0 ? a.a_bytes : a.b_bytes, kRsrcWord3);
  const int lrow = lane / KCh::kChunks, pc = lane % KCh::kChunks;
  // the swizzle of a load row depends on the load index only through (q * kRowsPerLoad) bits
  // above the swizzle's: KC = 16 never, KC = 32 through q's parity
  uint32_t lbyte[2];
#pragma unroll
  for (int par = 0; par < 2; ++par) {
    const int row = par * KCh::kRowsPerLoad + lrow;
    lbyte[par] = (uint32_t)(lrow * ld * 4 + ((pc ^ KCh::swz(row)) << 4));
  }
  const uint32_t rbytes = (uint32_t)(KCh::kRowsPerLoad * ld * 4);
  const uint32_t sbase = (uint32_t)(prow0 * ld * 4);
  char* const dst0 = smem + panel * KCh::kPanelBytes;


  const int sa = KCh::swz(ty), sb = KCh::swz(tx);
  const int64_t nk_all = a.Kp / KC;
  const int64_t kc0 = part < 0 ? 0 : part * a.kc_per;
  const int64_t nk = part < 0 ? nk_all : (kc0 + a.kc_per < nk_all ? kc0 + a.kc_per : nk_all);
  if (kc0 < nk)
    key_dma_issue<KC>(rsrc, lbyte, sbase + (uint32_t)(kc0 * KCh::kRowBytes), rbytes, nload, q0,
                      dst0 + (kc0 & 1) * KCh::kStageBytes);
  for (int64_t kc = kc0; kc < nk; ++kc) {
    // this wave's loads of chunk kc have landed; after the barrier, everyone's have, and every
    // wave is done reading the other stage (chunk kc - 1), which the next loads overwrite
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
    if (kc + 1 < nk)
      key_dma_issue<KC>(rsrc, lbyte, sbase + (uint32_t)((kc + 1) * KCh::kRowBytes), rbytes, nload, q0,
                        dst0 + ((kc + 1) & 1) * KCh::kStageBytes);
    const char* As = smem + (kc & 1) * KCh::kStageBytes;
    const char* Bs = same ? As : As + KCh::kPanelBytes;
#pragma unroll
    for (int c = 0; c < KCh::kChunks; ++c) {
      v4i av[8], bv[8];
      const int oa = ty * KCh::kRowBytes + ((c ^ sa) << 4);
      const int ob = tx * KCh::kRowBytes + ((c ^ sb) << 4);
#pragma unroll
      for (int m = 0; m < 8; ++m) av[m] = *reinterpret_cast<const v4i*>(As + oa + 16 * m * KCh::kRowBytes);
#pragma unroll
      for (int n = 0; n < 8; ++n) bv[n] = *reinterpret_cast<const v4i*>(Bs + ob + 16 * n * KCh::kRowBytes);
#pragma unroll
      for (int m = 0; m < 8; ++m)
#pragma unroll
        for (int n = 0; n < 8; ++n) acc[m][n] = step4(acc[m][n], av[m], bv[n]);
    }
  }

  if (part >= 0) {   // one K range of a split tail tile: fold into the tile's key block
    int* T = a.tail + (tile - a.nfull) * (int64_t)(kBM * kBN);
#pragma unroll
    for (int m = 0; m < 8; ++m)
#pragma unroll
      for (int n = 0; n < 8; ++n) atomicMax(T + (ty + 16 * m) * kBN + tx + 16 * n, acc[m][n]);
    return;
  }

  // epilogue: float and/or key outputs, optional fused ratio, mirror for symmetric off-diagonal tiles
  const bool mirror = a.tri && bm != bn;
#pragma unroll
  for (int m = 0; m < 8; ++m) {
    const int64_t i = row0 + ty + 16 * m;
    if (i >= a.M) continue;
#pragma unroll
    for (int n = 0; n < 8; ++n) {
      const int64_t j = col0 + tx + 16 * n;
      if (j >= a.N) continue;
      const float v = key_f32(acc[m][n]);
      if (!WIDE && a.C) {
        a.C[i * a.ldc + j] = v;
        if (mirror) a.C[j * a.ldc + i] = v;
      }
      if (a.Ck) {
        a.Ck[i * a.ldck + j] = acc[m][n];
        if (mirror) a.Ck[j * a.ldck + i] = acc[m][n];
      }
      if (!WIDE && a.O) {
        a.O[i * a.ldo + j] = ratio_of<float>(a.R[i * a.ldr + j], v);
        if (mirror) a.O[j * a.ldo + i] = ratio_of<float>(a.R[j * a.ldr + i], v);
      }
      if (WIDE) {
        const double w = key_f64(acc[m][n]);
        if (a.Cd) {
          a.Cd[i * a.ldc + j] = w;
          if (mirror) a.Cd[j * a.ldc + i] = w;
        }
        if (a.Od) {
          a.Od[i * a.ldo + j] = ratio_of<double>(a.Rd[i * a.ldr + j], w);
          if (mirror) a.Od[j * a.ldo + i] = ratio_of<double>(a.Rd[j * a.ldr + i], w);
        }
      }
    }
  }
}

// The outputs of the split tail tiles from their folded key blocks, as the key kernel's epilogue
// writes them.
__global__ void mm_finish_tail_kernel(MKArgs a) {
  const int64_t ntail = a.ntiles - a.nfull;
  const int64_t total = ntail * (kBM * kBN);
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t tt = t / (kBM * kBN), e = t % (kBM * kBN);
    int64_t bm, bn;
    mm_tile_coords(a, a.nfull + tt, bm, bn);
    const int64_t i = bm * kBM + e / kBN, j = bn * kBN + e % kBN;
    if (i >= a.M || j >= a.N) continue;
    const bool mirror = a.tri && bm != bn;
    const int k = a.tail[t];
    const float v = key_f32(k);
    if (a.C) {
      a.C[i * a.ldc + j] = v;
      if (mirror) a.C[j * a.ldc + i] = v;
    }
    if (a.Ck) {
      a.Ck[i * a.ldck + j] = k;
      if (mirror) a.Ck[j * a.ldck + i] = k;
    }
    if (a.O) {
      a.O[i * a.ldo + j] = ratio_of<float>(a.R[i * a.ldr + j], v);
      if (mirror) a.O[j * a.ldo + i] = ratio_of<float>(a.R[j * a.ldr + i], v);
    }
    if (a.Cd || a.Od) {
      const double w = key_f64(k);
      if (a.Cd) {
        a.Cd[i * a.ldc + j] = w;
        if (mirror) a.Cd[j * a.ldc + i] = w;
      }
      if (a.Od) {
        a.Od[i * a.ldo + j] = ratio_of<double>(a.Rd[i * a.ldr + j], w);
        if (mirror) a.Od[j * a.ldo + i] = ratio_of<double>(a.Rd[j * a.ldr + i], w);
      }
    }
  }
}

__global__ void fill_i32_kernel(int* p, int64_t n, int v) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) p[t] = v;
}

// ---------------------------------------------------------------- FP64 padded-value path
// FP64 operands are copied once into padded value matrices (rows padded to the tile, row length
// a multiple of the largest K chunk, padding -inf, the K mask folded in), so that K chunks arrive
// in LDS by direct buffer loads exactly as the FP32 keys do.  v_min_f64 / v_max_f64 are written
// as single instructions: as library calls the compiler quiets every LDS operand first
// (v_max_f64 x, x, x), an eighth more VALU work.  Split-tail partial maxima meet in a block of
// 64-bit order keys through atomicMax.
constexpr int kValAlign = 16;   // row length of value matrices (doubles): a multiple of the largest K chunk

__device__ __forceinline__ double vmin_f64(double a, double b) {
  double r;
  asm("v_min_f64 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
  return r;
}
__device__ __forceinline__ double vmax_f64(double a, double b) {
  double r;
  asm("v_max_f64 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
  return r;
}
__device__ __forceinline__ long long f64_key(double x) {
  const long long b = __double_as_longlong(x);
  return b ^ ((b >> 63) & 0x7fffffffffffffffll);
}
__device__ __forceinline__ double key_f64(long long k) {
  return __longlong_as_double(k ^ ((k >> 63) & 0x7fffffffffffffffll));
}

__global__ void val_prep_kernel(const double* __restrict__ X, int64_t ldx, int64_t rows, int64_t K,
                                const double* __restrict__ kmask, double* __restrict__ out, int64_t ld, int64_t rows_pad) {
  const int64_t total = rows_pad * ld;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = t / ld, k = t % ld;
    double v = -INFINITY;
    if (r < rows && k < K) {
      v = X[r * ldx + k];
      if (kmask) v = vmin_f64(v, kmask[k]);
    }
    out[t] = v;
  }
}

// -inf into the padding of a value matrix whose [0, rows) x [0, cols) part a product writes
__global__ void val_pad_kernel(double* __restrict__ out, int64_t ld, int64_t rows, int64_t cols, int64_t rows_pad) {
  const int64_t pad_cols = ld - cols;
  const int64_t total = rows * pad_cols + (rows_pad - rows) * ld;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
    if (t < rows * pad_cols) out[(t / pad_cols) * ld + cols + t % pad_cols] = -INFINITY;
    else out[rows * ld + (t - rows * pad_cols)] = -INFINITY;
  }
}

__global__ void fill_i64_kernel(long long* p, int64_t n, long long v) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) p[t] = v;
}

struct MVArgs {
  const double* A; int64_t lda;   // padded values, rows padded to kBM, lda a multiple of the K chunk
  const double* B; int64_t ldb;
  uint32_t a_bytes, b_bytes;      // buffer ranges
  int64_t M, N, Kp;               // Kp: padded K (doubles)
  double* C; int64_t ldc;         // output (nullable)
  double* Cp; int64_t ldcp;       // second output, e.g. a padded operand of the next product (nullable)
  const double* R; int64_t ldr;   // ratio numerator (nullable: no fused ratio)
  double* O; int64_t ldo;         // ratio output
  int tri, shared;
  int64_t nbm, nbn, ntiles;
  const uint32_t* order;          // tri mode: tile coordinates in launch order
  int64_t nfull;                  // split tail, as for the FP32 keys
  int splits;
  int64_t kc_per;
  long long* tail;
};

__device__ __forceinline__ void mv_tile_coords(const MVArgs& a, int64_t t, int64_t& bm, int64_t& bn) {
  if (a.tri) {
    const uint32_t code = a.order[t];
    bm = code >> 16;
    bn = code & 0xffffu;
  } else {
    const int64_t group = 8;
    const int64_t per_group = group * a.nbn;
    const int64_t g = t / per_group;
    const int64_t first = g * group;
    const int64_t gsize = (a.nbm - first) < group ? (a.nbm - first) : group;
    const int64_t in = t % per_group;
    bm = first + in % gsize;
    bn = in / gsize;
  }
}

__device__ __forceinline__ void mv_store(const MVArgs& a, int64_t i, int64_t j, bool mirror, double v) {
  if (a.C) {
    a.C[i * a.ldc + j] = v;
    if (mirror) a.C[j * a.ldc + i] = v;
  }
  if (a.Cp) {
    a.Cp[i * a.ldcp + j] = v;
    if (mirror) a.Cp[j * a.ldcp + i] = v;
  }
  if (a.O) {
    a.O[i * a.ldo + j] = ratio_of<double>(a.R[i * a.ldr + j], v);
    if (mirror) a.O[j * a.ldo + i] = ratio_of<double>(a.R[j * a.ldr + i], v);
  }
}

typedef double v2d __attribute__((ext_vector_type(2)));

// acc[m][n] = max(acc[m][n], max(min(a[m].x, b.x), min(a[m].y, b.y))) for m < 8, as one asm
// block in an order where no instruction reads the result of the one before.  As separate
// statements the compiler, which cannot see their latency, puts dependent ones next to each
// other and a hazard wait state between them (a third as many s_nop as min/max instructions).
template <typename Acc>
__device__ __forceinline__ void mm_f64_column(Acc& acc, int n, const v2d (&a)[8], const v2d& b) {
  double t0, t1, t2, t3;
  asm("v_min_f64 %[t0], %[a0x], %[bx]\n\t"
      "v_min_f64 %[t1], %[a0y], %[by]\n\t"
      "v_min_f64 %[t2], %[a1x], %[bx]\n\t"
      "v_min_f64 %[t3], %[a1y], %[by]\n\t"
      "v_max_f64 %[t0], %[t0], %[t1]\n\t"
      "v_max_f64 %[t2], %[t2], %[t3]\n\t"
      "v_max_f64 %[c0], %[c0], %[t0]\n\t"
      "v_max_f64 %[c1], %[c1], %[t2]\n\t"
      "v_min_f64 %[t0], %[a2x], %[bx]\n\t"
      "v_min_f64 %[t1], %[a2y], %[by]\n\t"
      "v_min_f64 %[t2], %[a3x], %[bx]\n\t"
      "v_min_f64 %[t3], %[a3y], %[by]\n\t"
      "v_max_f64 %[t0], %[t0], %[t1]\n\t"
      "v_max_f64 %[t2], %[t2], %[t3]\n\t"
      "v_max_f64 %[c2], %[c2], %[t0]\n\t"
      "v_max_f64 %[c3], %[c3], %[t2]\n\t"
      "v_min_f64 %[t0], %[a4x], %[bx]\n\t"
      "v_min_f64 %[t1], %[a4y], %[by]\n\t"
      "v_min_f64 %[t2], %[a5x], %[bx]\n\t"
      "v_min_f64 %[t3], %[a5y], %[by]\n\t"
      "v_max_f64 %[t0], %[t0], %[t1]\n\t"
      "v_max_f64 %[t2], %[t2], %[t3]\n\t"
      "v_max_f64 %[c4], %[c4], %[t0]\n\t"
      "v_max_f64 %[c5], %[c5], %[t2]\n\t"
      "v_min_f64 %[t0], %[a6x], %[bx]\n\t"
      "v_min_f64 %[t1], %[a6y], %[by]\n\t"
      "v_min_f64 %[t2], %[a7x], %[bx]\n\t"
      "v_min_f64 %[t3], %[a7y], %[by]\n\t"
      "v_max_f64 %[t0], %[t0], %[t1]\n\t"
      "v_max_f64 %[t2], %[t2], %[t3]\n\t"
      "v_max_f64 %[c6], %[c6], %[t0]\n\t"
      "v_max_f64 %[c7], %[c7], %[t2]"
      : [t0] "=&v"(t0), [t1] "=&v"(t1), [t2] "=&v"(t2), [t3] "=&v"(t3),
        [c0] "+v"(acc[0][n]), [c1] "+v"(acc[1][n]), [c2] "+v"(acc[2][n]), [c3] "+v"(acc[3][n]),
        [c4] "+v"(acc[4][n]), [c5] "+v"(acc[5][n]), [c6] "+v"(acc[6][n]), [c7] "+v"(acc[7][n])
      : [a0x] "v"(a[0].x), [a0y] "v"(a[0].y), [a1x] "v"(a[1].x), [a1y] "v"(a[1].y),
        [a2x] "v"(a[2].x), [a2y] "v"(a[2].y), [a3x] "v"(a[3].x), [a3y] "v"(a[3].y),
        [a4x] "v"(a[4].x), [a4y] "v"(a[4].y), [a5x] "v"(a[5].x), [a5y] "v"(a[5].y),
        [a6x] "v"(a[6].x), [a6y] "v"(a[6].y), [a7x] "v"(a[7].x), [a7y] "v"(a[7].y),
        [bx] "v"(b.x), [by] "v"(b.y));
}

// KD doubles of K per stage: LDS panel rows of KD * 8 bytes, with the swizzle of the FP32 key
// rows of the same length (KeyChunk<2 KD>).
template <int KD>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(2))) void maxmin_f64_kernel(MVArgs a) {
  using KCh = KeyChunk<2 * KD>;
  extern __shared__ __align__(16) char smem[];

  int64_t tile = blockIdx.x, part = -1;
  if (tile >= a.nfull) {
    part = (tile - a.nfull) % a.splits;
    tile = a.nfull + (tile - a.nfull) / a.splits;
  }
  int64_t bm, bn;
  mv_tile_coords(a, tile, bm, bn);
  const int64_t row0 = bm * kBM, col0 = bn * kBN;
  const bool same = a.shared && bm == bn;

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int tx = tid & 15, ty = tid >> 4;

  double acc[8][8];
#pragma unroll
  for (int m = 0; m < 8; ++m)
#pragma unroll
    for (int n = 0; n < 8; ++n) acc[m][n] = -DBL_MAX;   // numeric_limits<double>::lowest()

  const int swave = __builtin_amdgcn_readfirstlane(wave);
  const int panel = same ? 0 : (swave >> 1);
  const int nload = same ? KCh::kLoadsPerPanel / 4 : KCh::kLoadsPerPanel / 2;
  const int q0 = same ? swave * nload : (swave & 1) * nload;
  const int64_t prow0 = panel == 0 ? row0 : col0;
  const int64_t ld = panel == 0 ? a.lda : a.ldb;
  const __amdgpu_buffer_rsrc_t rsrc = __builtin_amdgcn_make_buffer_rsrc(
      (void*)(panel == 0 ? a.A : a.B), (short)0, panel == 0 ? a.a_bytes : a.b_bytes, kRsrcWord3);
  const int lrow = lane / KCh::kChunks, pc = lane % KCh::kChunks;
  uint32_t lbyte[2];
#pragma unroll
  for (int par = 0; par < 2; ++par) {
    const int row = par * KCh::kRowsPerLoad + lrow;
    lbyte[par] = (uint32_t)(lrow * ld * 8 + ((pc ^ KCh::swz(row)) << 4));
  }
  const uint32_t rbytes = (uint32_t)(KCh::kRowsPerLoad * ld * 8);
  const uint32_t sbase = (uint32_t)(prow0 * ld * 8);
  char* const dst0 = smem + panel * KCh::kPanelBytes;

  const int sa = KCh::swz(ty), sb = KCh::swz(tx);
  const int64_t nk_all = a.Kp / KD;
  const int64_t kc0 = part < 0 ? 0 : part * a.kc_per;
  const int64_t nk = part < 0 ? nk_all : (kc0 + a.kc_per < nk_all ? kc0 + a.kc_per : nk_all);
  if (kc0 < nk)
    key_dma_issue<2 * KD>(rsrc, lbyte, sbase + (uint32_t)(kc0 * KCh::kRowBytes), rbytes, nload, q0,
                          dst0 + (kc0 & 1) * KCh::kStageBytes);
  for (int64_t kc = kc0; kc < nk; ++kc) {
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
    if (kc + 1 < nk)
      key_dma_issue<2 * KD>(rsrc, lbyte, sbase + (uint32_t)((kc + 1) * KCh::kRowBytes), rbytes, nload, q0,
                            dst0 + ((kc + 1) & 1) * KCh::kStageBytes);
    const char* As = smem + (kc & 1) * KCh::kStageBytes;
    const char* Bs = same ? As : As + KCh::kPanelBytes;
#pragma unroll
    for (int c = 0; c < KCh::kChunks; ++c) {
      // the A fragments of the chunk stay in registers; each B fragment is read one use ahead
      v2d av[8];
      const int oa = ty * KCh::kRowBytes + ((c ^ sa) << 4);
      const int ob = tx * KCh::kRowBytes + ((c ^ sb) << 4);
#pragma unroll
      for (int m = 0; m < 8; ++m) av[m] = *reinterpret_cast<const v2d*>(As + oa + 16 * m * KCh::kRowBytes);
      v2d bnext = *reinterpret_cast<const v2d*>(Bs + ob);
#pragma unroll
      for (int n = 0; n < 8; ++n) {
        const v2d bv = bnext;
        if (n + 1 < 8) bnext = *reinterpret_cast<const v2d*>(Bs + ob + 16 * (n + 1) * KCh::kRowBytes);
        mm_f64_column(acc, n, av, bv);
      }
    }
  }

  if (part >= 0) {   // one K range of a split tail tile: fold into the tile's key block
    long long* T = a.tail + (tile - a.nfull) * (int64_t)(kBM * kBN);
#pragma unroll
    for (int m = 0; m < 8; ++m)
#pragma unroll
      for (int n = 0; n < 8; ++n) atomicMax(T + (ty + 16 * m) * kBN + tx + 16 * n, f64_key(acc[m][n]));
    return;
  }

  const bool mirror = a.tri && bm != bn;
#pragma unroll
  for (int m = 0; m < 8; ++m) {
    const int64_t i = row0 + ty + 16 * m;
    if (i >= a.M) continue;
#pragma unroll
    for (int n = 0; n < 8; ++n) {
      const int64_t j = col0 + tx + 16 * n;
      if (j >= a.N) continue;
      mv_store(a, i, j, mirror, acc[m][n]);
    }
  }
}

__global__ void mv_finish_tail_kernel(MVArgs a) {
  const int64_t ntail = a.ntiles - a.nfull;
  const int64_t total = ntail * (kBM * kBN);
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t tt = t / (kBM * kBN), e = t % (kBM * kBN);
    int64_t bm, bn;
    mv_tile_coords(a, a.nfull + tt, bm, bn);
    const int64_t i = bm * kBM + e / kBN, j = bn * kBN + e % kBN;
    if (i >= a.M || j >= a.N) continue;
    mv_store(a, i, j, a.tri && bm != bn, key_f64(a.tail[t]));
  }
}

template <typename T>
__global__ void ratio_kernel(const T* __restrict__ A, int64_t lda, const T* __restrict__ Mm, int64_t ldm,
                             int64_t M, int64_t N, T* __restrict__ O, int64_t ldo) {
  const int64_t total = M * N;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = t / N, j = t % N;
    O[i * ldo + j] = ratio_of<T>(A[i * lda + j], Mm[i * ldm + j]);
  }
}

template <typename T>
__global__ void emax_kernel(const T* __restrict__ A, int64_t lda, const T* __restrict__ B, int64_t ldb, int64_t M,
                            int64_t N, T* __restrict__ O, int64_t ldo) {
  const int64_t total = M * N;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = t / N, j = t % N;
    const T a = A[i * lda + j], b = B[i * ldb + j];
    O[i * ldo + j] = a < b ? b : a;   // std::max(a, b) returns a unless a < b
  }
}

// flags[0] |= 1 if A is not bitwise symmetric.  One workgroup per 64 x 64 tile (I, J) with
// I <= J (2-D grid; the others leave at once): tile (I, J) goes to LDS from coalesced row reads,
// tile (J, I) is read coalesced and compared with it transposed, so each element is read once.
constexpr int kSymTile = 64;

template <typename T, typename Bits = typename std::conditional<sizeof(T) == 4, uint32_t, uint64_t>::type>
__global__ __launch_bounds__(256) void symmetry_check_kernel(const T* __restrict__ A, int64_t lda, int64_t N,
                                                              int* flags) {
  if (blockIdx.y > blockIdx.x) return;
  __shared__ Bits tile[kSymTile][kSymTile + 1];
  const int64_t i0 = (int64_t)blockIdx.y * kSymTile, j0 = (int64_t)blockIdx.x * kSymTile;
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  for (int r = ty; r < kSymTile; r += 4) {
    const int64_t i = i0 + r, j = j0 + tx;
    tile[r][tx] = (i < N && j < N) ? __builtin_bit_cast(Bits, A[i * lda + j]) : Bits(0);
  }
  __syncthreads();
  int bad = 0;
  for (int r = ty; r < kSymTile; r += 4) {
    const int64_t j = j0 + r, i = i0 + tx;   // A[j][i] against A[i][j] = tile[tx][r]
    if (i < N && j < N && __builtin_bit_cast(Bits, A[j * lda + i]) != tile[tx][r]) bad = 1;
  }
  if (__syncthreads_or(bad) && threadIdx.x == 0) atomicOr(flags, 1);
}

template <typename T>
void launch_symmetry_check(mcpg_ctx* h, const T* A, int64_t lda, int64_t N) {
  const unsigned nb = (unsigned)ceil_div(N, kSymTile);
  hipLaunchKernelGGL(symmetry_check_kernel<T>, dim3(nb, nb), dim3(256), 0, h->stream, A, lda, N, h->d_flags);
}

// DPI decision and mask application (include/mcp/filter/dpi.hpp:56-66, :86-105;
// include/mcp/filter/threshold.hpp:34-54 with negate = true; apps/filter/dpi.cpp:236-285).  An edge
// x-z is removed when x == z or some allowed third gene y has min(A[x][y], A[z][y]) >= i_xz
// (EQUALITY_EXCLUSION), i.e. when W[x][z] = max over allowed y of that minimum reaches i_xz.
// i_xz = A[x][z] * tol_multiplier: a double product stored in T for dpi_kernel, a T product for
// dpi_tf_kernel (whose multiplier is stored in T).  With a TF list, pairs with exactly one TF
// endpoint use WT (y restricted to TFs).  Removed edges become 0 in O; `mask` (nullable) receives
// the decision itself (1 = remove), the boolean matrix InnerProduct<..., dpi_kernel> produces;
// `removed` counts the entries of O that changed.
template <typename T>
__global__ void dpi_apply_kernel(const T* __restrict__ A, int64_t lda, int64_t N, const T* __restrict__ W, int64_t ldw,
                                 const T* __restrict__ WT, int64_t ldwt, const T* __restrict__ tfs, double tm_d, T tm_t,
                                 T* __restrict__ O, int64_t ldo, uint8_t* __restrict__ mask, int64_t ldmask,
                                 unsigned long long* removed) {
  const int64_t total = N * N;
  unsigned long long cnt = 0;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t x = t / N, z = t % N;
    const T a = A[x * lda + z];
    bool del = true;
    if (x != z) {
      if (tfs) {
        const T ixz = a * tm_t;
        const bool tf_xz = (tfs[x] > (T)0) != (tfs[z] > (T)0);
        del = (tf_xz ? WT[x * ldwt + z] : W[x * ldw + z]) >= ixz;
      } else {
        const T ixz = (T)((double)a * tm_d);
        del = W[x * ldw + z] >= ixz;
      }
    }
    const T o = del ? (T)0 : a;
    if (O) O[x * ldo + z] = o;
    if (mask) mask[x * ldmask + z] = del ? 1 : 0;
    cnt += (o != a) ? 1 : 0;
  }
  for (int off = 32; off > 0; off >>= 1) cnt += __shfl_xor(cnt, off);
  if ((threadIdx.x & 63) == 0 && cnt) atomicAdd(removed, cnt);
}

// copy of A with the diagonal set to `diag` (FP64 DPI operand: y = x and y = z drop out of the max)
template <typename T>
__global__ void diag_copy_kernel(const T* __restrict__ A, int64_t lda, int64_t N, T diag, T* __restrict__ O) {
  const int64_t total = N * N;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = t / N, j = t % N;
    O[t] = i == j ? diag : A[i * lda + j];
  }
}

template <class Tr>
int launch_maxmin(mcpg_ctx* h, const typename Tr::T* A, int64_t lda, const typename Tr::T* B, int64_t ldb,
                  const typename Tr::T* kmask, int64_t M, int64_t N, int64_t K, int tri,
                  typename Tr::T* C, int64_t ldc, const typename Tr::T* R, int64_t ldr, typename Tr::T* O,
                  int64_t ldo) {
  using T = typename Tr::T;
  if (M <= 0 || N <= 0) return MCPG_OK;
  MMArgs a;
  a.A = A; a.lda = lda; a.B = B; a.ldb = ldb; a.kmask = kmask;
  a.M = M; a.N = N; a.K = K; a.C = C; a.ldc = ldc; a.R = R; a.ldr = ldr; a.O = O; a.ldo = ldo;
  a.tri = tri;
  a.shared = (A == B && lda == ldb && kmask == nullptr) ? 1 : 0;
  const bool al = (reinterpret_cast<uintptr_t>(A) % 16 == 0) && (reinterpret_cast<uintptr_t>(B) % 16 == 0) &&
                  ((lda * (int64_t)sizeof(T)) % 16 == 0) && ((ldb * (int64_t)sizeof(T)) % 16 == 0);
  a.aligned = al ? 1 : 0;
  a.nbm = ceil_div(M, kBM);
  a.nbn = ceil_div(N, kBN);
  a.ntiles = tri ? a.nbm * (a.nbm + 1) / 2 : a.nbm * a.nbn;
  if (a.ntiles > (int64_t)INT32_MAX) return mcpg::set_error(h, MCPG_ERR_UNSUPPORTED, "max-min grid too large");
  a.order = nullptr;
  if (tri) {
    int st;
    a.order = mcpg::tri_order(h, a.nbm, mcpg::env_group("MCPG_MM_GROUP", 8), mcpg::env_group("MCPG_MM_XCD", 0), &st);
    if (!a.order) return st;
  }
  {
    mcpg::LaunchScope ls(h, tri ? "maxmin_tile_kernel<f64,tri>" : "maxmin_tile_kernel<f64>");
    hipLaunchKernelGGL(maxmin_tile_kernel<Tr>, dim3((unsigned)a.ntiles), dim3(kThreads), 0, h->stream, a);
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return mcpg::hip_error(h, e, "maxmin_tile_kernel launch");
  return MCPG_OK;
}

// ---------------------------------------------------------------- FP32 key path, host side
constexpr int kKeyAlign = 32;   // row length of key matrices: a multiple of the largest K chunk

struct KeyMat {
  int* d = nullptr;
  int64_t rows = 0, rows_pad = 0, ldk = 0;
  uint32_t bytes() const { return (uint32_t)(rows_pad * ldk * 4); }
};

KeyMat key_geometry(int64_t rows, int64_t K) {
  KeyMat km;
  km.rows = rows;
  km.rows_pad = ceil_div(rows, kBM) * kBM;
  km.ldk = std::max<int64_t>(ceil_div(K, kKeyAlign) * kKeyAlign, kKeyAlign);
  return km;
}

bool key_fits(const KeyMat& km) { return (uint64_t)km.rows_pad * (uint64_t)km.ldk * 4 < (1ull << 32); }

int key_check(mcpg_ctx* h, const KeyMat& km) {
  if (!key_fits(km)) return mcpg::set_error(h, MCPG_ERR_UNSUPPORTED, "max-min operand exceeds 4 GiB of keys");
  return MCPG_OK;
}

int grid_of(int64_t total) { return (int)std::min<int64_t>(std::max<int64_t>(ceil_div(total, 256), 1), 65536); }

int prep_keys(mcpg_ctx* h, const float* X, int64_t ldx, int64_t K, const float* kmask, const KeyMat& km) {
  mcpg::LaunchScope ls(h, "key_prep_kernel");
  hipLaunchKernelGGL(key_prep_kernel, dim3(grid_of(km.rows_pad * km.ldk)), dim3(256), 0, h->stream, X, ldx, km.rows, K,
                     kmask, km.d, km.ldk, km.rows_pad);
  return MCPG_OK;
}

// padding of a key matrix whose [0, rows) x [0, cols) part a product will write
int pad_keys(mcpg_ctx* h, const KeyMat& km, int64_t cols) {
  const int64_t total = km.rows * (km.ldk - cols) + (km.rows_pad - km.rows) * km.ldk;
  if (total == 0) return MCPG_OK;
  mcpg::LaunchScope ls(h, "key_pad_kernel");
  hipLaunchKernelGGL(key_pad_kernel, dim3(grid_of(total)), dim3(256), 0, h->stream, km.d, km.ldk, km.rows, cols,
                     km.rows_pad);
  return MCPG_OK;
}

// Buffers a caller provides in place of the handle's cached tile orders and tail block (the
// allocation-free entry point): room for one triangle tile order and for split-tail key blocks.
// With too little tail room the product runs unsplit, which gives the same bits.
struct MMScratch {
  uint32_t* order = nullptr;
  int64_t order_cap = 0;   // entries
  int* tail = nullptr;
  size_t tail_bytes = 0;
};

// The order mcpg::tri_order builds without the XCD deal (blocks of group x group tiles in
// row-major order, tiles row-major inside a block, upper triangle only), written on the device
// by one thread: the allocation-free path has no cached table.
__global__ void tri_order_kernel(uint32_t* __restrict__ out, int64_t nb, int group) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  int64_t n = 0;
  for (int64_t bi = 0; bi < nb; bi += group)
    for (int64_t bj = 0; bj < nb; bj += group)
      for (int64_t ti = bi; ti < bi + group && ti < nb; ++ti)
        for (int64_t tj = bj > ti ? bj : ti; tj < bj + group && tj < nb; ++tj) out[n++] = (uint32_t)((ti << 16) | tj);
}

// FP64 outputs of a key product (see MKArgs::Cd): the max-min matrix and/or the ratio R / C
struct WideOut {
  double* C = nullptr;
  int64_t ldc = 0;
  const double* R = nullptr;
  int64_t ldr = 0;
  double* O = nullptr;
  int64_t ldo = 0;
};

// C / Ck / O over [0, M) x [0, N) of A (x) B, with K = A.ldk (== B.ldk); with `wide`, its double
// outputs in place of C and O.  With `sc`, the tile order and the tail blocks live in the
// caller's buffers and nothing is allocated.
int launch_maxmin_keys(mcpg_ctx* h, const KeyMat& A, const KeyMat& B, int64_t M, int64_t N, int tri, float* C,
                       int64_t ldc, int* Ck, int64_t ldck, const float* R, int64_t ldr, float* O, int64_t ldo,
                       const MMScratch* sc = nullptr, const WideOut* wide = nullptr, int64_t b0 = 0,
                       int64_t b1 = -1) {
  if (M <= 0 || N <= 0) return MCPG_OK;
  MKArgs a;
  a.A = A.d; a.lda = A.ldk; a.B = B.d; a.ldb = B.ldk; a.a_bytes = A.bytes(); a.b_bytes = B.bytes();
  a.M = M; a.N = N; a.Kp = A.ldk;
  a.C = C; a.ldc = ldc; a.Ck = Ck; a.ldck = ldck; a.R = R; a.ldr = ldr; a.O = O; a.ldo = ldo;
  a.Cd = nullptr; a.Rd = nullptr; a.Od = nullptr;
  if (wide) {
    if (C || O) return mcpg::set_error(h, MCPG_ERR_INVALID_ARG, "a key product writes float or double outputs, not both");
    a.Cd = wide->C; a.ldc = wide->ldc; a.Rd = wide->R; a.ldr = wide->ldr; a.Od = wide->O; a.ldo = wide->ldo;
  }
  a.tri = tri;
  a.shared = A.d == B.d ? 1 : 0;
  a.nbm = ceil_div(M, kBM);
  a.nbn = ceil_div(N, kBN);
  a.ntiles = tri ? a.nbm * (a.nbm + 1) / 2 : a.nbm * a.nbn;
  if (a.ntiles > (int64_t)INT32_MAX) return mcpg::set_error(h, MCPG_ERR_UNSUPPORTED, "max-min grid too large");
  // Row panel [b0, b1) of a triangle launch: the tiles with b0 <= bm < b1, a contiguous stretch of
  // the blocked order when b0 and b1 fall on its row groups (and tiles are not dealt over XCDs).
  // After the panels in increasing order the rows of every panel so far are complete.
  const bool panel = b0 != 0 || (b1 >= 0 && b1 != a.nbm);
  if (panel) {
    const int64_t group = std::max(mcpg::env_group("MCPG_MM_GROUP", 8), 1);
    if (!tri || sc || mcpg::env_group("MCPG_MM_XCD", 0) != 0 || b0 < 0 || b1 <= b0 || b1 > a.nbm ||
        b0 % group != 0 || (b1 % group != 0 && b1 != a.nbm))
      return mcpg::set_error(h, MCPG_ERR_INVALID_ARG, "bad row panel of a triangle max-min launch");
    a.ntiles = (b1 * a.nbm - b1 * (b1 - 1) / 2) - (b0 * a.nbm - b0 * (b0 - 1) / 2);
  }
  a.order = nullptr;
  if (tri && sc) {
    if (a.ntiles > sc->order_cap || a.nbm > 65535)
      return mcpg::set_error(h, MCPG_ERR_UNSUPPORTED, "max-min tile order exceeds the scratch buffer");
    const int group = std::max(mcpg::env_group("MCPG_MM_GROUP", 8), 1);
    mcpg::LaunchScope ls(h, "tri_order_kernel");
    hipLaunchKernelGGL(tri_order_kernel, dim3(1), dim3(64), 0, h->stream, sc->order, a.nbm, group);
    a.order = sc->order;
  } else if (tri) {
    int st;
    a.order = mcpg::tri_order(h, a.nbm, mcpg::env_group("MCPG_MM_GROUP", 8), mcpg::env_group("MCPG_MM_XCD", 0), &st);
    if (!a.order) return st;
    if (panel) a.order += b0 * a.nbm - b0 * (b0 - 1) / 2;
  }
  const int kc = mcpg::env_group("MCPG_MM_KC", 16) == 32 ? 32 : 16;
  const void* kfns[2][2] = {{reinterpret_cast<const void*>(&maxmin_key_kernel<16, false>),
                             reinterpret_cast<const void*>(&maxmin_key_kernel<16, true>)},
                            {reinterpret_cast<const void*>(&maxmin_key_kernel<32, false>),
                             reinterpret_cast<const void*>(&maxmin_key_kernel<32, true>)}};
  const void* kfn = kfns[kc == 16 ? 0 : 1][wide ? 1 : 0];
  const int smem = 2 * (kc == 16 ? KeyChunk<16>::kStageBytes : KeyChunk<32>::kStageBytes);
  static bool attr_set = false;
  if (!attr_set) {
    for (int c = 0; c < 2; ++c)
      for (int w = 0; w < 2; ++w)
        MCPG_HIP(h, hipFuncSetAttribute(kfns[c][w], hipFuncAttributeMaxDynamicSharedMemorySize,
                                        2 * (c == 0 ? KeyChunk<16>::kStageBytes : KeyChunk<32>::kStageBytes)));
    attr_set = true;
  }
  // Split tail: the tiles beyond the last full round of resident workgroups (all tiles, for a
  // grid smaller than one round) would leave most of the GPU idle for a whole tile time; they
  // run instead as several workgroups each over a range of K, folding exact partial maxima with
  // atomicMax on the keys.
  int& slots = h->mm_slots[(kc == 16 ? 0 : 1) + (wide ? 2 : 0)];
  if (slots == 0) {
    int per_cu = 0, cus = 0;
    MCPG_HIP(h, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kfn, kThreads, smem));
    MCPG_HIP(h, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, h->device));
    slots = std::max(1, per_cu * cus);
  }
  const int64_t nk = a.Kp / kc;
  a.nfull = a.ntiles;
  a.splits = 1;
  a.kc_per = nk;
  a.tail = nullptr;
  {
    const int64_t full = a.ntiles / slots * slots, ntail = a.ntiles - full;
    const int64_t splits = ntail > 0 ? std::min<int64_t>(slots / ntail, nk) : 1;
    const size_t need = (size_t)ntail * kBM * kBN * sizeof(int);
    if (mcpg::env_group("MCPG_MM_SPLIT", 1) != 0 && splits >= 2 && (!sc || need <= sc->tail_bytes)) {
      if (sc) {
        a.tail = sc->tail;
      } else if (need > h->mm_tail_bytes) {
        MCPG_HIP(h, hipStreamSynchronize(h->stream));
        if (h->mm_tail) (void)hipFree(h->mm_tail);
        h->mm_tail = nullptr;
        h->mm_tail_bytes = 0;
        MCPG_HIP(h, hipMalloc(&h->mm_tail, need));
        h->mm_tail_bytes = need;
      }
      a.nfull = full;
      a.splits = (int)splits;
      a.kc_per = ceil_div(nk, splits);
      if (!sc) a.tail = h->mm_tail;
      mcpg::LaunchScope ls(h, "fill_i32_kernel");
      hipLaunchKernelGGL(fill_i32_kernel, dim3(grid_of(ntail * kBM * kBN)), dim3(256), 0, h->stream, a.tail,
                         ntail * kBM * kBN, (int)INT32_MIN);
    }
  }
  const int64_t grid = a.nfull + (a.ntiles - a.nfull) * a.splits;
  if (grid > (int64_t)INT32_MAX) return mcpg::set_error(h, MCPG_ERR_UNSUPPORTED, "max-min grid too large");
  {
    mcpg::LaunchScope ls(h, wide ? (tri ? "maxmin_key_kernel<f64 out,tri>" : "maxmin_key_kernel<f64 out>")
                                 : (tri ? "maxmin_key_kernel<f32,tri>" : "maxmin_key_kernel<f32>"));
    void* args[] = {&a};
    const hipError_t le = hipLaunchKernel(kfn, dim3((unsigned)grid), dim3(kThreads), args, smem, h->stream);
    if (le != hipSuccess) return mcpg::hip_error(h, le, "maxmin_key_kernel launch");
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return mcpg::hip_error(h, e, "maxmin_key_kernel launch");
  if (a.nfull < a.ntiles) {
    mcpg::LaunchScope ls(h, "mm_finish_tail_kernel");
    hipLaunchKernelGGL(mm_finish_tail_kernel, dim3(grid_of((a.ntiles - a.nfull) * kBM * kBN)), dim3(256), 0, h->stream,
                       a);
    e = hipGetLastError();
    if (e != hipSuccess) return mcpg::hip_error(h, e, "mm_finish_tail_kernel launch");
  }
  return MCPG_OK;
}

// mcpg_maxmin_f32: keys of A (mask folded in) and of B (unless shared), then one product.  The key
// buffers are sized and carved by the caller together with its other workspace.
struct KeyPlan {
  KeyMat ka, kb;
  bool shareB = false;
};

// same_operand: A and B are the same matrix; a K mask applies to A only, so B then needs its own keys
KeyPlan plan_keys(bool same_operand, bool masked, int64_t M, int64_t N, int64_t K) {
  KeyPlan p;
  p.ka = key_geometry(M, K);
  p.shareB = same_operand && !masked;
  p.kb = p.shareB ? p.ka : key_geometry(N, K);
  return p;
}

void size_keys(mcpg::Sizer& sz, const KeyPlan& p) {
  sz.add<int>(p.ka.rows_pad * p.ka.ldk);
  if (!p.shareB) sz.add<int>(p.kb.rows_pad * p.kb.ldk);
}

void take_keys(mcpg::Carver& cv, KeyPlan& p) {
  p.ka.d = cv.take<int>(p.ka.rows_pad * p.ka.ldk);
  p.kb.d = p.shareB ? p.ka.d : cv.take<int>(p.kb.rows_pad * p.kb.ldk);
}

int run_keys(mcpg_ctx* h, const KeyPlan& p, const float* A, int64_t lda, const float* B, int64_t ldb, const float* kmask,
             int64_t K, int symmetric, float* C, int64_t ldc) {
  int st;
  if ((st = prep_keys(h, A, lda, K, kmask, p.ka))) return st;
  if (!p.shareB && (st = prep_keys(h, B, ldb, K, nullptr, p.kb))) return st;
  return launch_maxmin_keys(h, p.ka, p.kb, p.ka.rows, p.kb.rows, symmetric, C, ldc, nullptr, 0, nullptr, 0, nullptr, 0);
}

// ---- FP64 operands on the key kernel
// Max and min only select entries, so a product of FP64 operands whose entries a float holds
// exactly selects the same entries as the key product of their float images (key_prep_f64_kernel
// states the conditions); the key epilogue widens the result and takes the ratio in double.  The
// key kernel issues 1.5 VALU instructions per triple against the FP64 kernel's 2 (FP64 rate),
// and reads 4-byte keys.  MCPG_MM64_KEYS=0 keeps FP64 operands on the FP64 kernel.
bool f64_keys_enabled() { return mcpg::env_group("MCPG_MM64_KEYS", 1) != 0; }

int prep_keys(mcpg_ctx* h, const double* X, int64_t ldx, int64_t K, const double* kmask, const KeyMat& km) {
  mcpg::LaunchScope ls(h, "key_prep_f64_kernel");
  hipLaunchKernelGGL(key_prep_f64_kernel, dim3(grid_of(km.rows_pad * km.ldk)), dim3(256), 0, h->stream, X, ldx, km.rows,
                     K, kmask, km.d, km.ldk, km.rows_pad, h->d_flags);
  return MCPG_OK;
}

// after the FP64 operands' prep_keys calls (flags reset before them): whether every entry passed
int keys_exact(mcpg_ctx* h, bool* exact) {
  int flags[1];
  const int st = mcpg::read_flags(h, flags, 1);
  *exact = st == MCPG_OK && (flags[0] & 1) == 0;
  return st;
}

// One key product into outputs of the operands' type (see launch_maxmin_keys).
int key_product(mcpg_ctx* h, const KeyMat& A, const KeyMat& B, int64_t M, int64_t N, int tri, float* C, int64_t ldc,
                int* Ck, int64_t ldck, const float* R, int64_t ldr, float* O, int64_t ldo,
                const MMScratch* sc = nullptr, int64_t b0 = 0, int64_t b1 = -1) {
  return launch_maxmin_keys(h, A, B, M, N, tri, C, ldc, Ck, ldck, R, ldr, O, ldo, sc, nullptr, b0, b1);
}
int key_product(mcpg_ctx* h, const KeyMat& A, const KeyMat& B, int64_t M, int64_t N, int tri, double* C, int64_t ldc,
                int* Ck, int64_t ldck, const double* R, int64_t ldr, double* O, int64_t ldo,
                const MMScratch* sc = nullptr, int64_t b0 = 0, int64_t b1 = -1) {
  if (!C && !O)
    return launch_maxmin_keys(h, A, B, M, N, tri, nullptr, 0, Ck, ldck, nullptr, 0, nullptr, 0, sc, nullptr, b0, b1);
  WideOut w;
  w.C = C; w.ldc = ldc; w.R = R; w.ldr = ldr; w.O = O; w.ldo = ldo;
  return launch_maxmin_keys(h, A, B, M, N, tri, nullptr, 0, Ck, ldck, nullptr, 0, nullptr, 0, sc, &w, b0, b1);
}

// Row-block ends of the panels in which a hooked chain runs its last triangle product: 1/2, 3/4
// and 7/8 of the row groups, then the rest, so that each panel's rows (its download) fit under
// the next panel's product, whose share of the triangle shrinks with the row index.  One panel
// below four row groups, with tiles dealt over XCDs, or with MCPG_MM_PANELS=0.
std::vector<int64_t> row_panels(int64_t nb) {
  const int64_t group = std::max(mcpg::env_group("MCPG_MM_GROUP", 8), 1);
  const int64_t ng = ceil_div(nb, group);
  std::vector<int64_t> ends;
  if (mcpg::env_group("MCPG_MM_PANELS", 1) != 0 && mcpg::env_group("MCPG_MM_XCD", 0) == 0 && ng >= 4)
    for (int64_t eighths : {4, 6, 7}) {
      const int64_t e = std::min(nb, ceil_div(ng * eighths, 8) * group);
      if (e < nb && (ends.empty() || e > ends.back())) ends.push_back(e);
    }
  ends.push_back(nb);
  return ends;
}

// ---------------------------------------------------------------- FP64 padded-value path, host side
struct ValMat {
  double* d = nullptr;
  int64_t rows = 0, rows_pad = 0, ld = 0;
  uint32_t bytes() const { return (uint32_t)(rows_pad * ld * 8); }
};

ValMat val_geometry(int64_t rows, int64_t K) {
  ValMat vm;
  vm.rows = rows;
  vm.rows_pad = ceil_div(rows, kBM) * kBM;
  vm.ld = std::max<int64_t>(ceil_div(K, kValAlign) * kValAlign, kValAlign);
  return vm;
}

int val_check(mcpg_ctx* h, const ValMat& vm) {
  if ((uint64_t)vm.rows_pad * (uint64_t)vm.ld * 8 >= (1ull << 32))
    return mcpg::set_error(h, MCPG_ERR_UNSUPPORTED, "FP64 max-min operand exceeds 4 GiB");
  return MCPG_OK;
}

int prep_vals(mcpg_ctx* h, const double* X, int64_t ldx, int64_t K, const double* kmask, const ValMat& vm) {
  mcpg::LaunchScope ls(h, "val_prep_kernel");
  hipLaunchKernelGGL(val_prep_kernel, dim3(grid_of(vm.rows_pad * vm.ld)), dim3(256), 0, h->stream, X, ldx, vm.rows, K,
                     kmask, vm.d, vm.ld, vm.rows_pad);
  return MCPG_OK;
}

int pad_vals(mcpg_ctx* h, const ValMat& vm, int64_t cols) {
  const int64_t total = vm.rows * (vm.ld - cols) + (vm.rows_pad - vm.rows) * vm.ld;
  if (total == 0) return MCPG_OK;
  mcpg::LaunchScope ls(h, "val_pad_kernel");
  hipLaunchKernelGGL(val_pad_kernel, dim3(grid_of(total)), dim3(256), 0, h->stream, vm.d, vm.ld, vm.rows, cols,
                     vm.rows_pad);
  return MCPG_OK;
}

// C / Cp / O over [0, M) x [0, N) of A (x) B on padded FP64 values, K = A.ld (== B.ld)
int launch_maxmin_vals(mcpg_ctx* h, const ValMat& A, const ValMat& B, int64_t M, int64_t N, int tri, double* C,
                       int64_t ldc, double* Cp, int64_t ldcp, const double* R, int64_t ldr, double* O, int64_t ldo) {
  if (M <= 0 || N <= 0) return MCPG_OK;
  MVArgs a;
  a.A = A.d; a.lda = A.ld; a.B = B.d; a.ldb = B.ld; a.a_bytes = A.bytes(); a.b_bytes = B.bytes();
  a.M = M; a.N = N; a.Kp = A.ld;
  a.C = C; a.ldc = ldc; a.Cp = Cp; a.ldcp = ldcp; a.R = R; a.ldr = ldr; a.O = O; a.ldo = ldo;
  a.tri = tri;
  a.shared = A.d == B.d ? 1 : 0;
  a.nbm = ceil_div(M, kBM);
  a.nbn = ceil_div(N, kBN);
  a.ntiles = tri ? a.nbm * (a.nbm + 1) / 2 : a.nbm * a.nbn;
  if (a.ntiles > (int64_t)INT32_MAX) return mcpg::set_error(h, MCPG_ERR_UNSUPPORTED, "max-min grid too large");
  a.order = nullptr;
  if (tri) {
    int st;
    a.order = mcpg::tri_order(h, a.nbm, mcpg::env_group("MCPG_MM_GROUP", 8), mcpg::env_group("MCPG_MM_XCD", 0), &st);
    if (!a.order) return st;
  }
  const int kd = mcpg::env_group("MCPG_MM64_KD", 16) == 8 ? 8 : 16;
  const void* kfn = kd == 8 ? reinterpret_cast<const void*>(&maxmin_f64_kernel<8>)
                            : reinterpret_cast<const void*>(&maxmin_f64_kernel<16>);
  const int smem = 2 * (kd == 8 ? KeyChunk<16>::kStageBytes : KeyChunk<32>::kStageBytes);
  static bool attr_set = false;
  if (!attr_set) {
    MCPG_HIP(h, hipFuncSetAttribute(reinterpret_cast<const void*>(&maxmin_f64_kernel<8>),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, 2 * KeyChunk<16>::kStageBytes));
    MCPG_HIP(h, hipFuncSetAttribute(reinterpret_cast<const void*>(&maxmin_f64_kernel<16>),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, 2 * KeyChunk<32>::kStageBytes));
    attr_set = true;
  }
  int& slots = h->mm64_slots[kd == 8 ? 0 : 1];
  if (slots == 0) {
    int per_cu = 0, cus = 0;
    MCPG_HIP(h, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kfn, kThreads, smem));
    MCPG_HIP(h, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, h->device));
    slots = std::max(1, per_cu * cus);
  }
  const int64_t nk = a.Kp / kd;
  a.nfull = a.ntiles;
  a.splits = 1;
  a.kc_per = nk;
  a.tail = nullptr;
  {
    const int64_t full = a.ntiles / slots * slots, ntail = a.ntiles - full;
    const int64_t splits = ntail > 0 ? std::min<int64_t>(slots / ntail, nk) : 1;
    if (mcpg::env_group("MCPG_MM_SPLIT", 1) != 0 && splits >= 2) {
      const size_t need = (size_t)ntail * kBM * kBN * sizeof(long long);
      if (need > h->mm_tail_bytes) {
        MCPG_HIP(h, hipStreamSynchronize(h->stream));
        if (h->mm_tail) (void)hipFree(h->mm_tail);
        h->mm_tail = nullptr;
        h->mm_tail_bytes = 0;
        MCPG_HIP(h, hipMalloc(&h->mm_tail, need));
        h->mm_tail_bytes = need;
      }
      a.nfull = full;
      a.splits = (int)splits;
      a.kc_per = ceil_div(nk, splits);
      a.tail = reinterpret_cast<long long*>(h->mm_tail);
      mcpg::LaunchScope ls(h, "fill_i64_kernel");
      hipLaunchKernelGGL(fill_i64_kernel, dim3(grid_of(ntail * kBM * kBN)), dim3(256), 0, h->stream, a.tail,
                         ntail * kBM * kBN, (long long)LLONG_MIN);
    }
  }
  const int64_t grid = a.nfull + (a.ntiles - a.nfull) * a.splits;
  if (grid > (int64_t)INT32_MAX) return mcpg::set_error(h, MCPG_ERR_UNSUPPORTED, "max-min grid too large");
  {
    mcpg::LaunchScope ls(h, tri ? "maxmin_f64_kernel<tri>" : "maxmin_f64_kernel");
    if (kd == 8)
      hipLaunchKernelGGL(maxmin_f64_kernel<8>, dim3((unsigned)grid), dim3(kThreads), smem, h->stream, a);
    else
      hipLaunchKernelGGL(maxmin_f64_kernel<16>, dim3((unsigned)grid), dim3(kThreads), smem, h->stream, a);
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return mcpg::hip_error(h, e, "maxmin_f64_kernel launch");
  if (a.nfull < a.ntiles) {
    mcpg::LaunchScope ls(h, "mv_finish_tail_kernel");
    hipLaunchKernelGGL(mv_finish_tail_kernel, dim3(grid_of((a.ntiles - a.nfull) * kBM * kBN)), dim3(256), 0, h->stream,
                       a);
    e = hipGetLastError();
    if (e != hipSuccess) return mcpg::hip_error(h, e, "mv_finish_tail_kernel launch");
  }
  return MCPG_OK;
}

// The FP64 products run on padded values unless MCPG_MM64_TILE=1 selects the register-staged tile
// kernel (kept as a cross-check; the same bits).
bool f64_tile_kernel() { return mcpg::env_group("MCPG_MM64_TILE", 0) == 1; }

// C (and/or O) = A (x) B for raw FP64 operands through padded copies carved from `cv`
struct ValPlan {
  ValMat va, vb;
  bool shareB = false;
};

ValPlan plan_vals(bool same_operand, bool masked, int64_t M, int64_t N, int64_t K) {
  ValPlan p;
  p.va = val_geometry(M, K);
  p.shareB = same_operand && !masked;
  p.vb = p.shareB ? p.va : val_geometry(N, K);
  return p;
}

void size_vals(mcpg::Sizer& sz, const ValPlan& p) {
  sz.add<double>(p.va.rows_pad * p.va.ld);
  if (!p.shareB) sz.add<double>(p.vb.rows_pad * p.vb.ld);
}

void take_vals(mcpg::Carver& cv, ValPlan& p) {
  p.va.d = cv.take<double>(p.va.rows_pad * p.va.ld);
  p.vb.d = p.shareB ? p.va.d : cv.take<double>(p.vb.rows_pad * p.vb.ld);
}

int run_vals(mcpg_ctx* h, const ValPlan& p, const double* A, int64_t lda, const double* B, int64_t ldb,
             const double* kmask, int64_t K, int tri, double* C, int64_t ldc, const double* R, int64_t ldr, double* O,
             int64_t ldo) {
  int st;
  if ((st = prep_vals(h, A, lda, K, kmask, p.va))) return st;
  if (!p.shareB && (st = prep_vals(h, B, ldb, K, nullptr, p.vb))) return st;
  return launch_maxmin_vals(h, p.va, p.vb, p.va.rows, p.vb.rows, tri, C, ldc, nullptr, 0, R, ldr, O, ldo);
}

// compute_maxmins (apps/pipeline/mcpnet_helper.hpp:436-460) on FP32 keys: M1 = A (x) A,
// M2 = A (x) M1, M3 = M1 (x) M1, each written only when its output pointer is set.
// M1 = A (x) A, M2 = A (x) M1, M3 = M1 (x) M1 on keys.  FP64 operands: *done stays false when A
// fails the key test, and the caller runs the FP64 kernel.
template <typename T>
int maxmins_keys(mcpg_ctx* h, bool host, const T* A, int64_t lda, int64_t N, T* M1, int64_t ld1, T* M2, int64_t ld2,
                 T* M3, int64_t ld3, bool* done = nullptr, const mcpg::ChainHook* after = nullptr) {
  constexpr bool f64 = std::is_same<T, double>::value;
  const size_t NN = (size_t)N * N;
  KeyMat ka = key_geometry(N, N), k1 = key_geometry(N, N);
  int st;
  if (f64 && !key_fits(ka)) return MCPG_OK;   // the FP64 kernel decides
  if ((st = key_check(h, ka))) return st;
  mcpg::Sizer sz;
  if (host) { sz.add<T>(NN); for (T* p : {M1, M2, M3}) if (p) sz.add<T>(NN); }
  sz.add<int>(ka.rows_pad * ka.ldk);
  sz.add<int>(k1.rows_pad * k1.ldk);
  void* base = mcpg::workspace(h, sz.off, &st);
  if (!base) return st;
  mcpg::Carver cv(base);
  const T* dA = A;
  int64_t la = lda;
  T* dM[3] = {M1, M2, M3};
  int64_t ldM[3] = {ld1, ld2, ld3};
  if (host) {
    T* a = cv.take<T>(NN);
    if ((st = mcpg::copy_2d(h, a, N * sizeof(T), A, lda * sizeof(T), N * sizeof(T), N, hipMemcpyHostToDevice))) return st;
    dA = a; la = N;
    for (int q = 0; q < 3; ++q)
      if (dM[q]) { dM[q] = cv.take<T>(NN); ldM[q] = N; }
  }
  ka.d = cv.take<int>(ka.rows_pad * ka.ldk);
  k1.d = cv.take<int>(k1.rows_pad * k1.ldk);
  if (f64 && (st = mcpg::reset_flags(h))) return st;
  if ((st = prep_keys(h, dA, la, N, nullptr, ka))) return st;
  if constexpr (f64) {
    bool exact;
    if ((st = keys_exact(h, &exact)) || !exact) return st;
  }
  if ((st = pad_keys(h, k1, N))) return st;
  // Host outputs: each matrix (the last one row panel by row panel) goes down on the side stream
  // while the next product runs, one download after another (they share its pinned buffers).
  mcpg::Download dls[8];
  int ndl = 0;
  const mcpg::ChainHook down = [&](int q, int64_t r0, int64_t r1) -> int {
    T* const hout[3] = {M1, M2, M3};
    const int64_t ldh[3] = {ld1, ld2, ld3};
    int rc;
    if (!hout[q]) return MCPG_OK;
    if (ndl == 8) return mcpg::set_error(h, MCPG_ERR_UNSUPPORTED, "too many max-min downloads");
    if (ndl > 0 && (rc = dls[ndl - 1].wait(h))) return rc;
    return dls[ndl++].start(h, hout[q] + r0 * ldh[q], ldh[q] * sizeof(T), dM[q] + r0 * N, N * sizeof(T),
                            N * sizeof(T), r1 - r0);
  };
  if (host) after = &down;
  // Product q; with a hook, the last product of a triangle runs in row panels, each handed on as
  // soon as it is on the stream.
  const int qlast = dM[2] ? 2 : dM[1] ? 1 : 0;
  auto product = [&](int q, const KeyMat& x, const KeyMat& y, int tri, int* ck, int64_t ldck) -> int {
    const int64_t nb = ceil_div(N, kBM);
    const std::vector<int64_t> ends = after && tri && q == qlast ? row_panels(nb) : std::vector<int64_t>{nb};
    int64_t b0 = 0;
    for (const int64_t b1 : ends) {
      int rc;
      if ((rc = key_product(h, x, y, N, N, tri, dM[q], ldM[q], ck, ldck, nullptr, 0, nullptr, 0, nullptr, b0, b1)))
        return rc;
      if (after && (rc = (*after)(q, b0 * kBM, std::min(b1 * kBM, N)))) return rc;
      b0 = b1;
    }
    return MCPG_OK;
  };
  if ((st = product(0, ka, ka, 1, k1.d, k1.ldk))) return st;
  if (dM[1]) {
    if ((st = mcpg::reset_flags(h))) return st;
    {
      mcpg::LaunchScope ls(h, "symmetry_check_kernel");
      launch_symmetry_check<T>(h, dA, la, N);
    }
    int flags[1];
    if ((st = mcpg::read_flags(h, flags, 1))) return st;
    if ((st = product(1, ka, k1, flags[0] == 0, nullptr, 0))) return st;
  }
  if (dM[2] && (st = product(2, k1, k1, 1, nullptr, 0))) return st;
  for (int q = 0; q < ndl; ++q)
    if ((st = dls[q].wait(h))) return st;
  if (done) *done = true;
  return mcpg::finish(h, "mcpg_maxmins");
}

// DPI filter of an N x N MI matrix.  FP32: the max-min products run on keys with the excluded
// entries at INT_MIN; FP64: on a copy with a -inf diagonal and a -inf / +max TF column mask.
template <typename T>
int dpi_entry(mcpg_ctx* h, int ptrs, const T* A, int64_t lda, int64_t N, T tolerance, const T* tfs, T* out, int64_t ldo,
              uint8_t* mask, int64_t ldmask, int64_t* removed) {
  MCPG_CHECK_HANDLE(h);
  mcpg::begin_call(h);
  if (N < 0) return mcpg::set_error(h, MCPG_ERR_SHAPE, "negative dimension");
  if (removed) *removed = 0;
  if (N == 0) return MCPG_OK;
  if (!A || (!out && !mask)) return mcpg::set_error(h, MCPG_ERR_INVALID_ARG, "null matrix pointer");
  if (lda < N || (out && ldo < N) || (mask && ldmask < N))
    return mcpg::set_error(h, MCPG_ERR_INVALID_ARG, "leading dimension too small");
  if (!(tolerance >= (T)0 && tolerance < (T)1)) return mcpg::set_error(h, MCPG_ERR_INVALID_ARG, "tolerance must be in [0, 1)");
  if (ptrs != MCPG_HOST && ptrs != MCPG_DEVICE) return mcpg::set_error(h, MCPG_ERR_INVALID_ARG, "bad ptrs kind");
  MCPG_HIP(h, hipSetDevice(h->device));
  constexpr bool f32 = std::is_same<T, float>::value;
  const bool host = ptrs == MCPG_HOST;
  const size_t NN = (size_t)N * N;
  KeyMat kx = key_geometry(N, N), kt = key_geometry(N, N);
  int st;
  if (f32 && (st = key_check(h, kx))) return st;
  const bool try_keys = !f32 && f64_keys_enabled() && key_fits(kx);   // FP64 on keys when float-exact
  mcpg::Sizer sz;
  if (host) { sz.add<T>(NN); if (out) sz.add<T>(NN); if (mask) sz.add<uint8_t>(NN); }
  if (tfs) sz.add<T>(N);
  sz.add<T>(NN);                                  // W
  if (tfs) sz.add<T>(NN);                         // WT
  const bool vals = !f32 && !f64_tile_kernel();
  ValMat vx = val_geometry(N, N), vt = val_geometry(N, N);
  if (vals && (st = val_check(h, vx))) return st;
  if (f32 || try_keys) { sz.add<int>(kx.rows_pad * kx.ldk); if (tfs) sz.add<int>(kt.rows_pad * kt.ldk); }
  if (!f32) { sz.add<T>(NN); if (tfs) sz.add<T>(N); }  // diagonal-excluded copy, TF column mask
  if (vals) { sz.add<double>(vx.rows_pad * vx.ld); if (tfs) sz.add<double>(vt.rows_pad * vt.ld); }
  sz.add<unsigned long long>(1);
  void* base = mcpg::workspace(h, sz.off, &st);
  if (!base) return st;
  mcpg::Carver cv(base);
  const T* dA = A;
  int64_t la = lda;
  T* dO = out;
  int64_t lo = ldo;
  uint8_t* dmk = mask;
  int64_t lmk = ldmask;
  if (host) {
    T* a = cv.take<T>(NN);
    if ((st = mcpg::copy_2d(h, a, N * sizeof(T), A, lda * sizeof(T), N * sizeof(T), N, hipMemcpyHostToDevice))) return st;
    dA = a; la = N;
    if (out) { dO = cv.take<T>(NN); lo = N; }
    if (mask) { dmk = cv.take<uint8_t>(NN); lmk = N; }
  }
  T* dtf = nullptr;
  if (tfs) {
    dtf = cv.take<T>(N);
    MCPG_HIP(h, hipMemcpyAsync(dtf, tfs, N * sizeof(T), host ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice, h->stream));
  }
  T* W = cv.take<T>(NN);
  T* WT = tfs ? cv.take<T>(NN) : nullptr;
  bool keyed = f32;
  if (f32 || try_keys) {
    kx.d = cv.take<int>(kx.rows_pad * kx.ldk);
    if (tfs) kt.d = cv.take<int>(kt.rows_pad * kt.ldk);
    int* flags = nullptr;
    if (!f32) {
      if ((st = mcpg::reset_flags(h))) return st;
      flags = h->d_flags;
    }
    {
      mcpg::LaunchScope ls(h, "dpi_key_kernel");
      hipLaunchKernelGGL(dpi_key_kernel<T>, dim3(grid_of(kx.rows_pad * kx.ldk)), dim3(256), 0, h->stream, dA, la, N,
                         (const T*)nullptr, kx.d, kx.ldk, kx.rows_pad, flags);
      if (tfs)
        hipLaunchKernelGGL(dpi_key_kernel<T>, dim3(grid_of(kt.rows_pad * kt.ldk)), dim3(256), 0, h->stream, dA, la, N,
                           dtf, kt.d, kt.ldk, kt.rows_pad, flags);
    }
    if (!f32 && (st = keys_exact(h, &keyed))) return st;
  }
  if (keyed) {
    if ((st = key_product(h, kx, kx, N, N, 1, W, N, nullptr, 0, nullptr, 0, nullptr, 0))) return st;
    // max over TF y of min(A[x][y], A[z][y]): the mask on one operand suffices (min is associative)
    if (tfs && (st = key_product(h, kt, kx, N, N, 1, WT, N, nullptr, 0, nullptr, 0, nullptr, 0))) return st;
  } else {
    T* Ad = cv.take<T>(NN);
    {
      mcpg::LaunchScope ls(h, "diag_copy_kernel");
      hipLaunchKernelGGL(diag_copy_kernel<T>, dim3(grid_of(NN)), dim3(256), 0, h->stream, dA, la, N, (T)-INFINITY, Ad);
    }
    if (vals) {
      vx.d = cv.take<double>(vx.rows_pad * vx.ld);
      if ((st = prep_vals(h, (const double*)Ad, N, N, nullptr, vx))) return st;
      if ((st = launch_maxmin_vals(h, vx, vx, N, N, 1, (double*)W, N, nullptr, 0, nullptr, 0, nullptr, 0))) return st;
    } else if ((st = launch_maxmin<TrF64>(h, (const double*)Ad, N, (const double*)Ad, N, nullptr, N, N, N, 1,
                                          (double*)W, N, nullptr, 0, nullptr, 0))) {
      return st;
    }
    if (tfs) {
      std::vector<T> mask(N), htf(N);
      if (host) std::copy(tfs, tfs + N, htf.begin());
      else MCPG_HIP(h, hipMemcpy(htf.data(), tfs, N * sizeof(T), hipMemcpyDeviceToHost));
      for (int64_t k = 0; k < N; ++k) mask[k] = htf[k] > (T)0 ? std::numeric_limits<T>::max() : (T)-INFINITY;
      T* dmask = cv.take<T>(N);
      MCPG_HIP(h, hipMemcpy(dmask, mask.data(), N * sizeof(T), hipMemcpyHostToDevice));
      // max over TF y of min(A[x][y], A[z][y]): the mask on one operand suffices (min is associative)
      if (vals) {
        vt.d = cv.take<double>(vt.rows_pad * vt.ld);
        if ((st = prep_vals(h, (const double*)Ad, N, N, (const double*)dmask, vt))) return st;
        if ((st = launch_maxmin_vals(h, vt, vx, N, N, 1, (double*)WT, N, nullptr, 0, nullptr, 0, nullptr, 0))) return st;
      } else if ((st = launch_maxmin<TrF64>(h, (const double*)Ad, N, (const double*)Ad, N, (const double*)dmask, N, N,
                                            N, 1, (double*)WT, N, nullptr, 0, nullptr, 0))) {
        return st;
      }
    }
  }
  unsigned long long* dcount = cv.take<unsigned long long>(1);
  MCPG_HIP(h, hipMemsetAsync(dcount, 0, sizeof(unsigned long long), h->stream));
  const double tm_d = 1.0 / (1.0 - (double)tolerance);          // dpi_kernel: double multiplier
  const T tm_t = (T)(1.0 / (1.0 - (double)tolerance));          // dpi_tf_kernel: multiplier stored in T
  {
    mcpg::LaunchScope ls(h, "dpi_apply_kernel");
    hipLaunchKernelGGL(dpi_apply_kernel<T>, dim3(grid_of(NN)), dim3(256), 0, h->stream, dA, la, N, W, N, WT, N, dtf, tm_d,
                       tm_t, dO, lo, dmk, lmk, dcount);
  }
  unsigned long long cnt = 0;
  MCPG_HIP(h, hipMemcpyAsync(&cnt, dcount, sizeof cnt, hipMemcpyDeviceToHost, h->stream));
  if (host && out &&
      (st = mcpg::copy_2d(h, out, ldo * sizeof(T), dO, N * sizeof(T), N * sizeof(T), N, hipMemcpyDeviceToHost)))
    return st;
  if (host && mask && (st = mcpg::copy_2d(h, mask, ldmask, dmk, N, N, N, hipMemcpyDeviceToHost))) return st;
  if ((st = mcpg::finish(h, "mcpg_dpi"))) return st;
  if (removed) *removed = (int64_t)cnt;
  return MCPG_OK;
}

// compute_maxmins on padded FP64 values: A padded once, M1 written both to the caller and into a
// padded operand of the other two products
int maxmins_f64_vals(mcpg_ctx* h, bool host, const double* A, int64_t lda, int64_t N, double* M1, int64_t ld1,
                     double* M2, int64_t ld2, double* M3, int64_t ld3, const mcpg::ChainHook* after) {
  const size_t NN = (size_t)N * N;
  ValMat va = val_geometry(N, N), v1 = val_geometry(N, N);
  int st;
  if ((st = val_check(h, va))) return st;
  mcpg::Sizer sz;
  if (host) { sz.add<double>(NN); for (double* p : {M1, M2, M3}) if (p) sz.add<double>(NN); }
  sz.add<double>(va.rows_pad * va.ld);
  sz.add<double>(v1.rows_pad * v1.ld);
  void* base = mcpg::workspace(h, sz.off, &st);
  if (!base) return st;
  mcpg::Carver cv(base);
  const double* dA = A;
  int64_t la = lda;
  double* dM[3] = {M1, M2, M3};
  int64_t ldM[3] = {ld1, ld2, ld3};
  if (host) {
    double* a = cv.take<double>(NN);
    if ((st = mcpg::copy_2d(h, a, N * 8, A, lda * 8, N * 8, N, hipMemcpyHostToDevice))) return st;
    dA = a; la = N;
    for (int q = 0; q < 3; ++q)
      if (dM[q]) { dM[q] = cv.take<double>(NN); ldM[q] = N; }
  }
  va.d = cv.take<double>(va.rows_pad * va.ld);
  v1.d = cv.take<double>(v1.rows_pad * v1.ld);
  if ((st = prep_vals(h, dA, la, N, nullptr, va))) return st;
  if ((st = pad_vals(h, v1, N))) return st;
  if ((st = launch_maxmin_vals(h, va, va, N, N, 1, dM[0], ldM[0], v1.d, v1.ld, nullptr, 0, nullptr, 0))) return st;
  if (after && (st = (*after)(0, 0, N))) return st;
  if (dM[1]) {
    if ((st = mcpg::reset_flags(h))) return st;
    {
      mcpg::LaunchScope ls(h, "symmetry_check_kernel");
      launch_symmetry_check<double>(h, dA, la, N);
    }
    int flags[1];
    if ((st = mcpg::read_flags(h, flags, 1))) return st;
    if ((st = launch_maxmin_vals(h, va, v1, N, N, flags[0] == 0, dM[1], ldM[1], nullptr, 0, nullptr, 0, nullptr, 0)))
      return st;
    if (after && (st = (*after)(1, 0, N))) return st;
  }
  if (dM[2]) {
    if ((st = launch_maxmin_vals(h, v1, v1, N, N, 1, dM[2], ldM[2], nullptr, 0, nullptr, 0, nullptr, 0))) return st;
    if (after && (st = (*after)(2, 0, N))) return st;
  }
  if (host) {
    double* out[3] = {M1, M2, M3};
    const int64_t ldo[3] = {ld1, ld2, ld3};
    for (int q = 0; q < 3; ++q)
      if (out[q] && (st = mcpg::copy_2d(h, out[q], ldo[q] * 8, dM[q], N * 8, N * 8, N, hipMemcpyDeviceToHost))) return st;
  }
  return mcpg::finish(h, "mcpg_maxmins");
}

// compute_maxmins on FP64 values with the register-staged tile kernel
int maxmins_f64(mcpg_ctx* h, bool host, const double* A, int64_t lda, int64_t N, double* M1, int64_t ld1, double* M2,
                int64_t ld2, double* M3, int64_t ld3, const mcpg::ChainHook* after = nullptr) {
  if (!f64_tile_kernel()) return maxmins_f64_vals(h, host, A, lda, N, M1, ld1, M2, ld2, M3, ld3, after);
  const size_t NN = (size_t)N * N;
  mcpg::Sizer sz;
  if (host) { sz.add<double>(NN); for (double* p : {M2, M3}) if (p) sz.add<double>(NN); }
  if (host || !M1) sz.add<double>(NN);   // M1 is always needed on the device
  int st;
  void* base = mcpg::workspace(h, sz.off, &st);
  if (!base) return st;
  mcpg::Carver cv(base);
  const double* dA = A;
  int64_t la = lda;
  double* dM[3] = {M1, M2, M3};
  int64_t ldM[3] = {ld1, ld2, ld3};
  if (host) {
    double* a = cv.take<double>(NN);
    if ((st = mcpg::copy_2d(h, a, N * 8, A, lda * 8, N * 8, N, hipMemcpyHostToDevice))) return st;
    dA = a; la = N;
    for (int q = 1; q < 3; ++q)
      if (dM[q]) { dM[q] = cv.take<double>(NN); ldM[q] = N; }
  }
  if (host || !M1) { dM[0] = cv.take<double>(NN); ldM[0] = N; }
  if ((st = launch_maxmin<TrF64>(h, dA, la, dA, la, nullptr, N, N, N, 1, dM[0], ldM[0], nullptr, 0, nullptr, 0))) return st;
  if (after && (st = (*after)(0, 0, N))) return st;
  if (dM[1]) {
    if ((st = mcpg::reset_flags(h))) return st;
    {
      mcpg::LaunchScope ls(h, "symmetry_check_kernel");
      launch_symmetry_check<double>(h, dA, la, N);
    }
    int flags[1];
    if ((st = mcpg::read_flags(h, flags, 1))) return st;
    if ((st = launch_maxmin<TrF64>(h, dA, la, dM[0], ldM[0], nullptr, N, N, N, flags[0] == 0, dM[1], ldM[1], nullptr, 0,
                                   nullptr, 0)))
      return st;
    if (after && (st = (*after)(1, 0, N))) return st;
  }
  if (dM[2]) {
    if ((st = launch_maxmin<TrF64>(h, dM[0], ldM[0], dM[0], ldM[0], nullptr, N, N, N, 1, dM[2], ldM[2], nullptr, 0,
                                   nullptr, 0)))
      return st;
    if (after && (st = (*after)(2, 0, N))) return st;
  }
  if (host) {
    double* out[3] = {M1, M2, M3};
    const int64_t ldo[3] = {ld1, ld2, ld3};
    for (int q = 0; q < 3; ++q)
      if (out[q] && (st = mcpg::copy_2d(h, out[q], ldo[q] * 8, dM[q], N * 8, N * 8, N, hipMemcpyDeviceToHost))) return st;
  }
  return mcpg::finish(h, "mcpg_maxmins");
}

template <typename T>
int maxmins_entry(mcpg_ctx* h, int ptrs, const T* A, int64_t lda, int64_t N, T* M1, int64_t ld1, T* M2, int64_t ld2,
                  T* M3, int64_t ld3, const mcpg::ChainHook* after = nullptr) {
  MCPG_CHECK_HANDLE(h);
  mcpg::begin_call(h);
  if (N < 0) return mcpg::set_error(h, MCPG_ERR_SHAPE, "negative dimension");
  if (N == 0) return MCPG_OK;
  if (!A) return mcpg::set_error(h, MCPG_ERR_INVALID_ARG, "null matrix pointer");
  if (lda < N || (M1 && ld1 < N) || (M2 && ld2 < N) || (M3 && ld3 < N))
    return mcpg::set_error(h, MCPG_ERR_INVALID_ARG, "leading dimension too small");
  if (ptrs != MCPG_HOST && ptrs != MCPG_DEVICE) return mcpg::set_error(h, MCPG_ERR_INVALID_ARG, "bad ptrs kind");
  MCPG_HIP(h, hipSetDevice(h->device));
  if constexpr (std::is_same<T, float>::value) {
    return maxmins_keys<float>(h, ptrs == MCPG_HOST, A, lda, N, M1, ld1, M2, ld2, M3, ld3, nullptr, after);
  } else {
    bool done = false;
    if (f64_keys_enabled()) {
      const int st = maxmins_keys<double>(h, ptrs == MCPG_HOST, A, lda, N, M1, ld1, M2, ld2, M3, ld3, &done, after);
      if (st || done) return st;
    }
    return maxmins_f64(h, ptrs == MCPG_HOST, A, lda, N, M1, ld1, M2, ld2, M3, ld3, after);
  }
}

template <typename T>
struct TraitsOf;
template <>
struct TraitsOf<float> { using type = TrF32; };
template <>
struct TraitsOf<double> { using type = TrF64; };

// mcpg_maxmin_f64 on the key kernel; *done stays false when an operand or the mask fails the key
// test (or K is 0), and the caller runs the FP64 kernel.
int maxmin_f64_keys(mcpg_ctx* h, bool host, const double* A, int64_t lda, const double* B, int64_t ldb,
                    const double* kmask, int64_t M, int64_t N, int64_t K, int symmetric, double* C, int64_t ldc,
                    bool* done) {
  if (K == 0) return MCPG_OK;
  const bool shareB = (A == B && M == N && lda == ldb);
  KeyPlan kp = plan_keys(shareB, kmask != nullptr, M, N, K);
  if (!key_fits(kp.ka) || !key_fits(kp.kb)) return MCPG_OK;
  mcpg::Sizer sz;
  if (host) {
    sz.add<double>(M * K);
    if (!shareB) sz.add<double>(N * K);
    if (kmask) sz.add<double>(K);
    sz.add<double>(M * N);
  }
  size_keys(sz, kp);
  int st;
  void* base = mcpg::workspace(h, sz.off, &st);
  if (!base) return st;
  mcpg::Carver cv(base);
  const double *a = A, *b = B, *m = kmask;
  int64_t la = lda, lb = ldb, lc = ldc;
  double* c = C;
  if (host) {
    double* dA = cv.take<double>(M * K);
    double* dB = shareB ? dA : cv.take<double>(N * K);
    double* dm = kmask ? cv.take<double>(K) : nullptr;
    c = cv.take<double>(M * N);
    if ((st = mcpg::copy_2d(h, dA, K * 8, A, lda * 8, K * 8, M, hipMemcpyHostToDevice))) return st;
    if (!shareB && (st = mcpg::copy_2d(h, dB, K * 8, B, ldb * 8, K * 8, N, hipMemcpyHostToDevice))) return st;
    if (kmask) MCPG_HIP(h, hipMemcpyAsync(dm, kmask, K * 8, hipMemcpyHostToDevice, h->stream));
    a = dA; b = dB; m = dm; la = lb = K; lc = N;
  }
  take_keys(cv, kp);
  if ((st = mcpg::reset_flags(h))) return st;
  if ((st = prep_keys(h, a, la, K, m, kp.ka))) return st;
  if (!kp.shareB && (st = prep_keys(h, b, lb, K, nullptr, kp.kb))) return st;
  bool exact;
  if ((st = keys_exact(h, &exact)) || !exact) return st;
  if ((st = key_product(h, kp.ka, kp.kb, M, N, symmetric, c, lc, nullptr, 0, nullptr, 0, nullptr, 0))) return st;
  if (host && (st = mcpg::copy_2d(h, C, ldc * 8, c, N * 8, N * 8, M, hipMemcpyDeviceToHost))) return st;
  *done = true;
  return mcpg::finish(h, "mcpg_maxmin");
}

template <typename T>
int maxmin_entry(mcpg_ctx* h, int ptrs, const T* A, int64_t lda, const T* B, int64_t ldb, const T* kmask,
                 int64_t M, int64_t N, int64_t K, int symmetric, T* C, int64_t ldc) {
  using Tr = typename TraitsOf<T>::type;
  MCPG_CHECK_HANDLE(h);
  mcpg::begin_call(h);
  if (M < 0 || N < 0 || K < 0) return mcpg::set_error(h, MCPG_ERR_SHAPE, "negative dimension");
  if (M == 0 || N == 0) return MCPG_OK;
  if (!A || !B || !C) return mcpg::set_error(h, MCPG_ERR_INVALID_ARG, "null matrix pointer");
  if (lda < K || ldb < K || ldc < N) return mcpg::set_error(h, MCPG_ERR_INVALID_ARG, "leading dimension too small");
  if (symmetric && (A != B || M != N || lda != ldb))
    return mcpg::set_error(h, MCPG_ERR_INVALID_ARG, "symmetric max-min needs A == B and M == N");
  if (ptrs != MCPG_HOST && ptrs != MCPG_DEVICE) return mcpg::set_error(h, MCPG_ERR_INVALID_ARG, "bad ptrs kind");
  MCPG_HIP(h, hipSetDevice(h->device));
  constexpr bool f32 = std::is_same<T, float>::value;
  const bool host = ptrs == MCPG_HOST;
  if constexpr (!f32) {
    bool done = false;
    if (f64_keys_enabled()) {
      const int st = maxmin_f64_keys(h, host, A, lda, B, ldb, kmask, M, N, K, symmetric, C, ldc, &done);
      if (st || done) return st;
    }
  }
  const bool shareB = (A == B && M == N && lda == ldb);
  // FP32 runs on order keys, FP64 on padded values (both sized here with the host staging)
  KeyPlan kp;
  ValPlan vp;
  const bool vals = !f32 && !f64_tile_kernel();
  int st;
  if constexpr (f32) {
    kp = plan_keys(shareB, kmask != nullptr, M, N, K);
    if ((st = key_check(h, kp.ka)) || (st = key_check(h, kp.kb))) return st;
  } else if (vals) {
    vp = plan_vals(shareB, kmask != nullptr, M, N, K);
    if ((st = val_check(h, vp.va)) || (st = val_check(h, vp.vb))) return st;
  }
  mcpg::Sizer sz;
  if (host) {
    sz.add<T>(M * K);
    if (!shareB) sz.add<T>(N * K);
    if (kmask) sz.add<T>(K);
    sz.add<T>(M * N);
  }
  if constexpr (f32) size_keys(sz, kp);
  if (vals) size_vals(sz, vp);
  void* base = sz.off ? mcpg::workspace(h, sz.off, &st) : h->ws;
  if (sz.off && !base) return st;
  mcpg::Carver cv(base);
  auto run = [&](const T* a, int64_t la, const T* b, int64_t lb, const T* m, T* c, int64_t lc) -> int {
    if constexpr (f32) return run_keys(h, kp, a, la, b, lb, m, K, symmetric, c, lc);
    else if (vals)
      return run_vals(h, vp, a, la, b, lb, m, K, symmetric, c, lc, nullptr, 0, nullptr, 0);
    else return launch_maxmin<Tr>(h, a, la, b, lb, m, M, N, K, symmetric, c, lc, nullptr, 0, nullptr, 0);
  };
  if (!host) {
    if constexpr (f32) take_keys(cv, kp);
    if (vals) take_vals(cv, vp);
    if ((st = run(A, lda, B, ldb, kmask, C, ldc))) return st;
    return mcpg::finish(h, "mcpg_maxmin");
  }
  T* dA = cv.take<T>(M * K);
  T* dB = shareB ? dA : cv.take<T>(N * K);
  T* dm = kmask ? cv.take<T>(K) : nullptr;
  T* dC = cv.take<T>(M * N);
  if constexpr (f32) take_keys(cv, kp);
  if (vals) take_vals(cv, vp);
  if (K > 0) {
    if ((st = mcpg::copy_2d(h, dA, K * sizeof(T), A, lda * sizeof(T), K * sizeof(T), M, hipMemcpyHostToDevice))) return st;
    if (!shareB && (st = mcpg::copy_2d(h, dB, K * sizeof(T), B, ldb * sizeof(T), K * sizeof(T), N, hipMemcpyHostToDevice))) return st;
    if (kmask) MCPG_HIP(h, hipMemcpyAsync(dm, kmask, K * sizeof(T), hipMemcpyHostToDevice, h->stream));
  }
  st = run(dA, K, dB, K, dm, dC, N);
  if (st) return st;
  if ((st = mcpg::copy_2d(h, C, ldc * sizeof(T), dC, N * sizeof(T), N * sizeof(T), M, hipMemcpyDeviceToHost))) return st;
  return mcpg::finish(h, "mcpg_maxmin");
}

template <typename T>
int ratio_entry(mcpg_ctx* h, int ptrs, const T* A, int64_t lda, const T* Mm, int64_t ldm, int64_t M, int64_t N,
                T* O, int64_t ldo, bool elementwise_max = false) {
  MCPG_CHECK_HANDLE(h);
  mcpg::begin_call(h);
  if (M < 0 || N < 0) return mcpg::set_error(h, MCPG_ERR_SHAPE, "negative dimension");
  if (M == 0 || N == 0) return MCPG_OK;
  if (!A || !Mm || !O) return mcpg::set_error(h, MCPG_ERR_INVALID_ARG, "null matrix pointer");
  if (lda < N || ldm < N || ldo < N) return mcpg::set_error(h, MCPG_ERR_INVALID_ARG, "leading dimension too small");
  MCPG_HIP(h, hipSetDevice(h->device));
  const T *dA = A, *dM = Mm;
  T* dO = O;
  int64_t la = lda, lm = ldm, lo = ldo;
  int st;
  if (ptrs == MCPG_HOST) {
    mcpg::Sizer sz;
    sz.add<T>(M * N); sz.add<T>(M * N); sz.add<T>(M * N);
    void* base = mcpg::workspace(h, sz.off, &st);
    if (!base) return st;
    mcpg::Carver cv(base);
    T* a = cv.take<T>(M * N); T* m = cv.take<T>(M * N); T* o = cv.take<T>(M * N);
    if ((st = mcpg::copy_2d(h, a, N * sizeof(T), A, lda * sizeof(T), N * sizeof(T), M, hipMemcpyHostToDevice))) return st;
    if ((st = mcpg::copy_2d(h, m, N * sizeof(T), Mm, ldm * sizeof(T), N * sizeof(T), M, hipMemcpyHostToDevice))) return st;
    dA = a; dM = m; dO = o; la = lm = lo = N;
  } else if (ptrs != MCPG_DEVICE) {
    return mcpg::set_error(h, MCPG_ERR_INVALID_ARG, "bad ptrs kind");
  }
  const int64_t total = M * N;
  const int blocks = (int)std::min<int64_t>(ceil_div(total, 256), 65536);
  if (elementwise_max) {
    mcpg::LaunchScope ls(h, "emax_kernel");
    hipLaunchKernelGGL(emax_kernel<T>, dim3(blocks), dim3(256), 0, h->stream, dA, la, dM, lm, M, N, dO, lo);
  } else {
    mcpg::LaunchScope ls(h, "ratio_kernel");
    hipLaunchKernelGGL(ratio_kernel<T>, dim3(blocks), dim3(256), 0, h->stream, dA, la, dM, lm, M, N, dO, lo);
  }
  if (ptrs == MCPG_HOST &&
      (st = mcpg::copy_2d(h, O, ldo * sizeof(T), dO, N * sizeof(T), N * sizeof(T), M, hipMemcpyDeviceToHost)))
    return st;
  return mcpg::finish(h, "mcpg_ratio");
}

// Static scratch of the allocation-free MCP entry point (mcpg_mcp_f32_noalloc).  Each device the
// library runs on holds its own copy until the process ends; host code reaches it only through
// the per-device address taken with hipGetSymbolAddress at handle creation (mcpg_ctx::scratch).
constexpr size_t kScratchBytes = size_t(64) << 20;
static __device__ __align__(256) char scratch[kScratchBytes];

// Scratch plan of mcp_keys<float> with device operands: keys of A, keys of M1 (k >= 3), one
// triangle tile order; the rest of `cap` bytes (if any) holds split-tail key blocks.
size_t mcp_scratch_bytes(int64_t N, int k) {
  const KeyMat ka = key_geometry(N, N);
  mcpg::Sizer sz;
  sz.add<int>(ka.rows_pad * ka.ldk);
  if (k >= 3) sz.add<int>(ka.rows_pad * ka.ldk);
  const int64_t nb = ceil_div(N, kBM);
  sz.add<uint32_t>(nb * (nb + 1) / 2);
  return sz.off;
}

// MCP2/3/4 on FP32 order keys: keys of A once; M1 = A (x) A stays in key form on the device; the
// last product writes the max-min matrix only when asked for and fuses the ratio.  With
// `scratch_base` (device operands only) every buffer is carved from those `scratch_cap` bytes and
// nothing is allocated.
template <typename T>
int mcp_keys(mcpg_ctx* h, bool host, const T* A, int64_t lda, int64_t N, int k, T* mm_out, int64_t ldm, T* out,
             int64_t ldo, void* scratch_base = nullptr, size_t scratch_cap = 0, bool* done = nullptr) {
  constexpr bool f64 = std::is_same<T, double>::value;
  const size_t NN = (size_t)N * N;
  KeyMat ka = key_geometry(N, N), k1 = key_geometry(N, N);
  int st;
  if (f64 && !key_fits(ka)) return MCPG_OK;   // the FP64 kernel decides
  if ((st = key_check(h, ka))) return st;
  mcpg::Sizer sz;
  if (host) { sz.add<T>(NN); sz.add<T>(NN); if (mm_out) sz.add<T>(NN); }
  sz.add<int>(ka.rows_pad * ka.ldk);
  if (k >= 3) sz.add<int>(k1.rows_pad * k1.ldk);
  MMScratch msc;
  const MMScratch* sc = nullptr;
  void* base = nullptr;
  if (scratch_base) {
    if (host) return mcpg::set_error(h, MCPG_ERR_INVALID_ARG, "the scratch path takes device operands only");
    const int64_t nb = ceil_div(N, kBM);
    const size_t keys = sz.off;
    sz.add<uint32_t>(nb * (nb + 1) / 2);
    if (sz.off > scratch_cap) return mcpg::set_error(h, MCPG_ERR_UNSUPPORTED, "MCP operands exceed the scratch buffer");
    base = scratch_base;
    msc.order = reinterpret_cast<uint32_t*>(static_cast<char*>(base) + keys);
    msc.order_cap = nb * (nb + 1) / 2;
    msc.tail = reinterpret_cast<int*>(static_cast<char*>(base) + sz.off);
    msc.tail_bytes = scratch_cap - sz.off;
    sc = &msc;
  } else {
    base = mcpg::workspace(h, sz.off, &st);
    if (!base) return st;
  }
  mcpg::Carver cv(base);
  const T* dA = A;
  int64_t la = lda;
  T* dO = out;
  int64_t lo = ldo;
  T* dM = mm_out;
  int64_t lm = ldm;
  if (host) {
    T* a = cv.take<T>(NN);
    if ((st = mcpg::copy_2d(h, a, N * sizeof(T), A, lda * sizeof(T), N * sizeof(T), N, hipMemcpyHostToDevice))) return st;
    dA = a; la = N;
    dO = cv.take<T>(NN); lo = N;
    if (mm_out) { dM = cv.take<T>(NN); lm = N; }
  }
  ka.d = cv.take<int>(ka.rows_pad * ka.ldk);
  if (f64 && (st = mcpg::reset_flags(h))) return st;
  if ((st = prep_keys(h, dA, la, N, nullptr, ka))) return st;
  if constexpr (f64) {
    bool exact;
    if ((st = keys_exact(h, &exact)) || !exact) return st;
  }
  // The last product (ratio fused).  Host outputs: a triangle runs in row panels, and each
  // panel's rows go down on the side stream (one download after another: they share its pinned
  // buffers) while the next panel runs; every download ends before the call returns.
  mcpg::Download dls[8];
  int ndl = 0;
  auto last = [&](const KeyMat& x, const KeyMat& y, int tri) -> int {
    const int64_t nb = ceil_div(N, kBM);
    const std::vector<int64_t> ends = host && !sc && tri ? row_panels(nb) : std::vector<int64_t>{nb};
    int64_t b0 = 0;
    int rc;
    for (const int64_t b1 : ends) {
      if ((rc = key_product(h, x, y, N, N, tri, dM, lm, nullptr, 0, dA, la, dO, lo, sc, b0, b1))) return rc;
      const int64_t r0 = b0 * kBM, rows = std::min(b1 * kBM, N) - r0;
      for (int o = 0; host && o < (mm_out ? 2 : 1); ++o) {
        if (ndl == 8) return mcpg::set_error(h, MCPG_ERR_UNSUPPORTED, "too many MCP downloads");
        if (ndl > 0 && (rc = dls[ndl - 1].wait(h))) return rc;
        T* dst = o ? mm_out : out;
        const int64_t ld = o ? ldm : ldo;
        const T* src = (o ? dM : dO) + r0 * N;
        if ((rc = dls[ndl++].start(h, dst + r0 * ld, ld * sizeof(T), src, N * sizeof(T), N * sizeof(T), rows)))
          return rc;
      }
      b0 = b1;
    }
    for (int q = 0; q < ndl; ++q)
      if ((rc = dls[q].wait(h))) return rc;
    return MCPG_OK;
  };
  if (k == 2) {
    if ((st = last(ka, ka, 1))) return st;
  } else {
    k1.d = cv.take<int>(k1.rows_pad * k1.ldk);
    if ((st = pad_keys(h, k1, N))) return st;
    // the first product A (x) A is symmetric by construction
    if ((st = key_product(h, ka, ka, N, N, 1, (T*)nullptr, 0, k1.d, k1.ldk, nullptr, 0, nullptr, 0, sc))) return st;
    if (k == 3) {
      // A (x) M1 is exactly symmetric when A is (max/min never round); check A bitwise.  The
      // scratch path clears the flag with a kernel (no memset node for callers that capture).
      if (sc) {
        mcpg::LaunchScope ls(h, "fill_i32_kernel");
        hipLaunchKernelGGL(fill_i32_kernel, dim3(1), dim3(64), 0, h->stream, h->d_flags, mcpg::kFlagWords, 0);
      } else if ((st = mcpg::reset_flags(h))) {
        return st;
      }
      {
        mcpg::LaunchScope ls(h, "symmetry_check_kernel");
        launch_symmetry_check<T>(h, dA, la, N);
      }
      int flags[1];
      if ((st = mcpg::read_flags(h, flags, 1))) return st;
      if ((st = last(ka, k1, flags[0] == 0))) return st;
    } else {
      if ((st = last(k1, k1, 1))) return st;
    }
  }
  if (done) *done = true;
  return mcpg::finish(h, "mcpg_mcp");
}

// MCP2/3/4 on padded FP64 values: A padded once; M1 = A (x) A is written straight into a padded
// operand for the second product; the last product writes the max-min matrix only when asked for
// and fuses the ratio.
int mcp_f64_vals(mcpg_ctx* h, bool host, const double* A, int64_t lda, int64_t N, int k, double* mm_out, int64_t ldm,
                 double* out, int64_t ldo) {
  const size_t NN = (size_t)N * N;
  ValMat va = val_geometry(N, N), v1 = val_geometry(N, N);
  int st;
  if ((st = val_check(h, va))) return st;
  mcpg::Sizer sz;
  if (host) { sz.add<double>(NN); sz.add<double>(NN); if (mm_out) sz.add<double>(NN); }
  sz.add<double>(va.rows_pad * va.ld);
  if (k >= 3) sz.add<double>(v1.rows_pad * v1.ld);
  void* base = mcpg::workspace(h, sz.off, &st);
  if (!base) return st;
  mcpg::Carver cv(base);
  const double* dA = A;
  int64_t la = lda;
  double* dO = out;
  int64_t lo = ldo;
  double* dM = mm_out;
  int64_t lm = ldm;
  if (host) {
    double* a = cv.take<double>(NN);
    if ((st = mcpg::copy_2d(h, a, N * 8, A, lda * 8, N * 8, N, hipMemcpyHostToDevice))) return st;
    dA = a; la = N;
    dO = cv.take<double>(NN); lo = N;
    if (mm_out) { dM = cv.take<double>(NN); lm = N; }
  }
  va.d = cv.take<double>(va.rows_pad * va.ld);
  if ((st = prep_vals(h, dA, la, N, nullptr, va))) return st;
  if (k == 2) {
    if ((st = launch_maxmin_vals(h, va, va, N, N, 1, dM, lm, nullptr, 0, dA, la, dO, lo))) return st;
  } else {
    v1.d = cv.take<double>(v1.rows_pad * v1.ld);
    if ((st = pad_vals(h, v1, N))) return st;
    // the first product A (x) A is symmetric by construction
    if ((st = launch_maxmin_vals(h, va, va, N, N, 1, nullptr, 0, v1.d, v1.ld, nullptr, 0, nullptr, 0))) return st;
    if (k == 3) {
      // A (x) M1 is exactly symmetric when A is (max/min never round); check A bitwise.
      if ((st = mcpg::reset_flags(h))) return st;
      {
        mcpg::LaunchScope ls(h, "symmetry_check_kernel");
        launch_symmetry_check<double>(h, dA, la, N);
      }
      int flags[1];
      if ((st = mcpg::read_flags(h, flags, 1))) return st;
      if ((st = launch_maxmin_vals(h, va, v1, N, N, flags[0] == 0, dM, lm, nullptr, 0, dA, la, dO, lo))) return st;
    } else {
      if ((st = launch_maxmin_vals(h, v1, v1, N, N, 1, dM, lm, nullptr, 0, dA, la, dO, lo))) return st;
    }
  }
  if (host) {
    if ((st = mcpg::copy_2d(h, out, ldo * 8, dO, N * 8, N * 8, N, hipMemcpyDeviceToHost))) return st;
    if (mm_out && (st = mcpg::copy_2d(h, mm_out, ldm * 8, dM, N * 8, N * 8, N, hipMemcpyDeviceToHost))) return st;
  }
  return mcpg::finish(h, "mcpg_mcp");
}

template <typename T>
int mcp_entry(mcpg_ctx* h, int ptrs, const T* A, int64_t lda, int64_t N, int k, T* mm_out, int64_t ldm, T* out,
              int64_t ldo) {
  using Tr = typename TraitsOf<T>::type;
  MCPG_CHECK_HANDLE(h);
  mcpg::begin_call(h);
  if (k < 2 || k > 4) return mcpg::set_error(h, MCPG_ERR_INVALID_ARG, "MCP order must be 2, 3 or 4");
  if (N < 0) return mcpg::set_error(h, MCPG_ERR_SHAPE, "negative dimension");
  if (N == 0) return MCPG_OK;
  if (!A || !out) return mcpg::set_error(h, MCPG_ERR_INVALID_ARG, "null matrix pointer");
  if (lda < N || ldo < N || (mm_out && ldm < N)) return mcpg::set_error(h, MCPG_ERR_INVALID_ARG, "leading dimension too small");
  if (ptrs != MCPG_HOST && ptrs != MCPG_DEVICE) return mcpg::set_error(h, MCPG_ERR_INVALID_ARG, "bad ptrs kind");
  MCPG_HIP(h, hipSetDevice(h->device));
  if constexpr (std::is_same<T, float>::value) {
    return mcp_keys<float>(h, ptrs == MCPG_HOST, A, lda, N, k, mm_out, ldm, out, ldo);
  } else {
    bool done = false;
    if (f64_keys_enabled()) {
      const int st = mcp_keys<double>(h, ptrs == MCPG_HOST, A, lda, N, k, mm_out, ldm, out, ldo, nullptr, 0, &done);
      if (st || done) return st;
    }
    if (!f64_tile_kernel()) return mcp_f64_vals(h, ptrs == MCPG_HOST, A, lda, N, k, mm_out, ldm, out, ldo);
  }
  const size_t NN = (size_t)N * N;
  const bool host = ptrs == MCPG_HOST;
  mcpg::Sizer sz;
  if (host) { sz.add<T>(NN); sz.add<T>(NN); }   // A copy, out
  sz.add<T>(NN);                                 // M1
  if (k >= 3 && (host || !mm_out)) sz.add<T>(NN);  // final max-min
  int st;
  void* base = mcpg::workspace(h, sz.off, &st);
  if (!base) return st;
  mcpg::Carver cv(base);
  const T* dA = A;
  int64_t la = lda;
  T* dO = out;
  int64_t lo = ldo;
  if (host) {
    T* a = cv.take<T>(NN);
    if ((st = mcpg::copy_2d(h, a, N * sizeof(T), A, lda * sizeof(T), N * sizeof(T), N, hipMemcpyHostToDevice))) return st;
    dA = a; la = N;
    dO = cv.take<T>(NN); lo = N;
  }
  T* M1 = cv.take<T>(NN);
  T* Mf = nullptr;  // final max-min matrix
  int64_t lf = N;
  if (k >= 3) {
    if (!host && mm_out) { Mf = mm_out; lf = ldm; }
    else Mf = cv.take<T>(NN);
  }
  // the first product A (x) A is symmetric by construction
  if (k == 2) {
    T* dst = (!host && mm_out) ? mm_out : M1;
    int64_t ld = (!host && mm_out) ? ldm : N;
    if ((st = launch_maxmin<Tr>(h, dA, la, dA, la, nullptr, N, N, N, 1, dst, ld, dA, la, dO, lo))) return st;
    Mf = dst; lf = ld;
  } else {
    if ((st = launch_maxmin<Tr>(h, dA, la, dA, la, nullptr, N, N, N, 1, M1, N, nullptr, 0, nullptr, 0))) return st;
    if (k == 3) {
      // A (x) M1 is exactly symmetric when A is (max/min never round); check A bitwise.
      if ((st = mcpg::reset_flags(h))) return st;
      {
        mcpg::LaunchScope ls(h, "symmetry_check_kernel");
        launch_symmetry_check<T>(h, dA, la, N);
      }
      int flags[1];
      if ((st = mcpg::read_flags(h, flags, 1))) return st;
      const int sym = flags[0] == 0;
      if ((st = launch_maxmin<Tr>(h, dA, la, M1, N, nullptr, N, N, N, sym, Mf, lf, dA, la, dO, lo))) return st;
    } else {
      if ((st = launch_maxmin<Tr>(h, M1, N, M1, N, nullptr, N, N, N, 1, Mf, lf, dA, la, dO, lo))) return st;
    }
  }
  if (host) {
    if ((st = mcpg::copy_2d(h, out, ldo * sizeof(T), dO, N * sizeof(T), N * sizeof(T), N, hipMemcpyDeviceToHost))) return st;
    if (mm_out && (st = mcpg::copy_2d(h, mm_out, ldm * sizeof(T), Mf, lf * sizeof(T), N * sizeof(T), N, hipMemcpyDeviceToHost)))
      return st;
  }
  return mcpg::finish(h, "mcpg_mcp");
}

// MCP2/3/4 of mcp_keys<float> on device operands with every buffer in `scratch`.
int mcp_noalloc_entry(mcpg_ctx* h, const float* A, int64_t lda, int64_t N, int k, float* mm_out, int64_t ldm, float* out,
                      int64_t ldo) {
  MCPG_CHECK_HANDLE(h);
  mcpg::begin_call(h);
  if (k < 2 || k > 4) return mcpg::set_error(h, MCPG_ERR_INVALID_ARG, "MCP order must be 2, 3 or 4");
  if (N < 0) return mcpg::set_error(h, MCPG_ERR_SHAPE, "negative dimension");
  if (N == 0) return MCPG_OK;
  if (!A || !out) return mcpg::set_error(h, MCPG_ERR_INVALID_ARG, "null matrix pointer");
  if (lda < N || ldo < N || (mm_out && ldm < N)) return mcpg::set_error(h, MCPG_ERR_INVALID_ARG, "leading dimension too small");
  if (!h->scratch)
    return mcpg::set_error(h, MCPG_ERR_HIP, std::string("static scratch unavailable: hipGetSymbolAddress: ") +
                                                hipGetErrorString(static_cast<hipError_t>(h->scratch_err)));
  if (mcp_scratch_bytes(N, k) > kScratchBytes) {
    char buf[160];
    snprintf(buf, sizeof buf, "MCP%d of %lld genes needs %zu bytes of scratch; the static scratch holds %zu", k,
             (long long)N, mcp_scratch_bytes(N, k), kScratchBytes);
    return mcpg::set_error(h, MCPG_ERR_UNSUPPORTED, buf);
  }
  MCPG_HIP(h, hipSetDevice(h->device));
  // launch timing would create events; this entry point records none
  const bool prof = h->profiling;
  h->profiling = false;
  const int st = mcp_keys<float>(h, false, A, lda, N, k, mm_out, ldm, out, ldo, h->scratch, kScratchBytes);
  h->profiling = prof;
  return st;
}

std::mutex g_scratch_mu;
constexpr int kMaxDevices = 64;
void* g_scratch_ptr[kMaxDevices] = {};

}  // namespace

namespace mcpg {

// Called by mcpg_create with the handle's device current: the device address of `scratch` on
// that device, taken once per device and kept for the process.
void init_scratch(mcpg_ctx* h) {
  h->scratch = nullptr;
  h->scratch_err = hipSuccess;
  if (h->device < 0 || h->device >= kMaxDevices) {
    h->scratch_err = hipErrorInvalidDevice;
    return;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[h->device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      (void)hipGetLastError();
      h->scratch_err = e != hipSuccess ? e : hipErrorInvalidSymbol;
      return;
    }
    g_scratch_ptr[h->device] = p;
  }
  h->scratch = g_scratch_ptr[h->device];
}

int maxmins_chain(mcpg_ctx* h, const float* A, int64_t N, float* M1, float* M2, float* M3,
                  const mcpg::ChainHook* after) {
  return maxmins_entry<float>(h, MCPG_DEVICE, A, N, N, M1, N, M2, N, M3, N, after);
}
int maxmins_chain(mcpg_ctx* h, const double* A, int64_t N, double* M1, double* M2, double* M3,
                  const mcpg::ChainHook* after) {
  return maxmins_entry<double>(h, MCPG_DEVICE, A, N, N, M1, N, M2, N, M3, N, after);
}

}  // namespace mcpg

extern "C" {

int mcpg_mcp_f32_noalloc(mcpg_handle h, const float* A, int64_t lda, int64_t N, int k, float* maxmin_out, int64_t ldm,
                         float* mcp_out, int64_t ldo) {
  return mcp_noalloc_entry(h, A, lda, N, k, maxmin_out, ldm, mcp_out, ldo);
}

int mcpg_maxmin_f32(mcpg_handle h, int ptrs, const float* A, int64_t lda, const float* B, int64_t ldb,
                    const float* kmask, int64_t M, int64_t N, int64_t K, int symmetric, float* C, int64_t ldc) {
  return maxmin_entry<float>(h, ptrs, A, lda, B, ldb, kmask, M, N, K, symmetric, C, ldc);
}
int mcpg_maxmin_f64(mcpg_handle h, int ptrs, const double* A, int64_t lda, const double* B, int64_t ldb,
                    const double* kmask, int64_t M, int64_t N, int64_t K, int symmetric, double* C, int64_t ldc) {
  return maxmin_entry<double>(h, ptrs, A, lda, B, ldb, kmask, M, N, K, symmetric, C, ldc);
}
int mcpg_ratio_f32(mcpg_handle h, int ptrs, const float* A, int64_t lda, const float* Mm, int64_t ldm, int64_t M,
                   int64_t N, float* out, int64_t ldo) {
  return ratio_entry<float>(h, ptrs, A, lda, Mm, ldm, M, N, out, ldo);
}
int mcpg_ratio_f64(mcpg_handle h, int ptrs, const double* A, int64_t lda, const double* Mm, int64_t ldm, int64_t M,
                   int64_t N, double* out, int64_t ldo) {
  return ratio_entry<double>(h, ptrs, A, lda, Mm, ldm, M, N, out, ldo);
}
int mcpg_elementwise_max_f32(mcpg_handle h, int ptrs, const float* A, int64_t lda, const float* B, int64_t ldb,
                             int64_t M, int64_t N, float* out, int64_t ldo) {
  return ratio_entry<float>(h, ptrs, A, lda, B, ldb, M, N, out, ldo, true);
}
int mcpg_elementwise_max_f64(mcpg_handle h, int ptrs, const double* A, int64_t lda, const double* B, int64_t ldb,
                             int64_t M, int64_t N, double* out, int64_t ldo) {
  return ratio_entry<double>(h, ptrs, A, lda, B, ldb, M, N, out, ldo, true);
}
int mcpg_mcp_f32(mcpg_handle h, int ptrs, const float* A, int64_t lda, int64_t N, int k, float* maxmin_out,
                 int64_t ldm, float* mcp_out, int64_t ldo) {
  return mcp_entry<float>(h, ptrs, A, lda, N, k, maxmin_out, ldm, mcp_out, ldo);
}
int mcpg_mcp_f64(mcpg_handle h, int ptrs, const double* A, int64_t lda, int64_t N, int k, double* maxmin_out,
                 int64_t ldm, double* mcp_out, int64_t ldo) {
  return mcp_entry<double>(h, ptrs, A, lda, N, k, maxmin_out, ldm, mcp_out, ldo);
}
int mcpg_maxmins_f32(mcpg_handle h, int ptrs, const float* A, int64_t lda, int64_t N, float* M1, int64_t ld1, float* M2,
                     int64_t ld2, float* M3, int64_t ld3) {
  return maxmins_entry<float>(h, ptrs, A, lda, N, M1, ld1, M2, ld2, M3, ld3);
}
int mcpg_maxmins_f64(mcpg_handle h, int ptrs, const double* A, int64_t lda, int64_t N, double* M1, int64_t ld1,
                     double* M2, int64_t ld2, double* M3, int64_t ld3) {
  return maxmins_entry<double>(h, ptrs, A, lda, N, M1, ld1, M2, ld2, M3, ld3);
}
int mcpg_dpi_f32(mcpg_handle h, int ptrs, const float* A, int64_t lda, int64_t N, float tolerance, const float* tfs,
                 float* out, int64_t ldo, uint8_t* mask, int64_t ldmask, int64_t* removed) {
  return dpi_entry<float>(h, ptrs, A, lda, N, tolerance, tfs, out, ldo, mask, ldmask, removed);
}
int mcpg_dpi_f64(mcpg_handle h, int ptrs, const double* A, int64_t lda, int64_t N, double tolerance, const double* tfs,
                 double* out, int64_t ldo, uint8_t* mask, int64_t ldmask, int64_t* removed) {
  return dpi_entry<double>(h, ptrs, A, lda, N, tolerance, tfs, out, ldo, mask, ldmask, removed);
}

}  // extern "C"
