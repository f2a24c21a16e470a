// B-spline mutual information: per-gene weight panels, the all-pairs MI contraction on MFMA
// with a fused p*log2(p) epilogue, and its adjoint.
//
// Layout in HBM.  Weights live in "group-tiled" row panels: rows are (gene, bin) pairs, samples
// run along the row (K of the contraction), zero-padded to Sp = roundup(S, 32).  Genes are
// packed GW = 80 / bins to an 80-row group (8 genes x 10 bins for the default), so a group is
// exactly five 16-row MFMA blocks; the remaining 80 - GW*bins rows of a group are zero.  Within
// a group the samples are blocked by 32 (Layout::at), so the slice a K step reads is one
// contiguous 80 x 32 run.  A workgroup tile is two groups (GT = 2*GW genes, 160 rows) per side.
//
// Arithmetic.  The joint table of genes (i, j) is the 10 x 10 block W_i * W_j^T / S of the
// (rows x rows) Gram matrix; it is never written to HBM.  Three contraction precisions share
// one tile loop: FP32 (v_mfma_f32_16x16x4_f32), BF16 (v_mfma_f32_16x16x32_bf16) and BF16X3,
// where each weight is split w = hi + lo into two BF16 values and the product is
// hi*hi + hi*lo + lo*hi with FP32 accumulation.  Marginal entropies are FP64 and
// MI = (H_i + H_j) + sum p log2 p is formed in FP64 from the FP32 entropy sum.
#include <cfloat>
#include <cmath>
#include <cstring>
#include <vector>

#include "internal.h"

#pragma clang fp contract(off)

namespace {

using mcpg::ceil_div;

constexpr int kMaxBins = 80;
constexpr int kMaxOrder = 8;
constexpr int kGroupRows = 80;       // rows of one gene group (5 MFMA blocks of 16)
constexpr int kTileRows = 160;       // two groups per workgroup tile side
constexpr int kKStep = 32;           // samples per K step (FP32 and BF16 planes)
constexpr int kKStep8 = 64;          // samples per K step of the int8 digit planes (64 B per row, as BF16)
constexpr int kThreads = 256;

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef int v4i __attribute__((ext_vector_type(4)));   // 16-byte staging chunk (a native vector, kept in registers)

struct Spline {
  int bins, order;
  float knots[kMaxBins + kMaxOrder];
};

// buildKnotVector<float> (BSplineMI.hpp:35-71): `order` zeros, bins-order internal knots
// i/(bins-order+1), `order` ones; values are products of an int and a double stored as float.
Spline make_spline(int bins, int order) {
  Spline sp;
  std::memset(&sp, 0, sizeof sp);
  sp.bins = bins;
  sp.order = order;
  const int n_int = bins - order;
  const double norm = 1.0 / (double)(n_int + 1);
  for (int i = 0; i < bins + order; ++i) {
    if (i < order) sp.knots[i] = 0.0f;
    else if (i < order + n_int) sp.knots[i] = (float)((double)(float)(i - order + 1) * norm);
    else sp.knots[i] = 1.0f;
  }
  return sp;
}

// Cox-de Boor basis values N_{b,order}(t) for b < bins, evaluated bottom-up with the same
// operations in T (float for the reference's float DataType, double for its double), degenerate-
// span guards (d < 1e-10) and negative truncation as the reference's recursive basisFunction
// (BSplineMI.hpp:130-164), including the right-end rule N_{bins-1,1}(t) = 1 when
// |t - knot[bins]| < 1e-10.  `lower` (nullable) receives the order-1 level N_{b,order-1}(t) for
// b < bins+1, used by the derivative.  NB / NO are compile-time capacities (bins <= NB,
// order <= NO) so that the tables stay in registers.
__device__ __forceinline__ double abs_of(float v) { return (double)fabsf(v); }
__device__ __forceinline__ double abs_of(double v) { return fabs(v); }

template <int NB, int NO, typename T>
__device__ __forceinline__ void basis_values_t(const T* k, int bins, int order, T t, T* out, T* lower) {
  T N[NB + NO];
  const int n1 = bins + order - 1;
#pragma unroll
  for (int i = 0; i < NB + NO - 1; ++i) {
    if (i < n1) {
      const bool in = (t >= k[i] && t < k[i + 1] && k[i] < k[i + 1]);
      const bool endpt = (abs_of(t - k[i + 1]) < 1e-10) && (i + 1 == bins);
      N[i] = (in || endpt) ? (T)1 : (T)0;
    } else {
      N[i] = (T)0;
    }
  }
  N[NB + NO - 1] = (T)0;
#pragma unroll
  for (int p = 2; p <= NO; ++p) {
    if (p <= order) {
      if (lower && p == order) {
#pragma unroll
        for (int i = 0; i < NB + 1; ++i) lower[i] = N[i];
      }
      const int n = bins + order - p;
#pragma unroll
      for (int i = 0; i < NB + NO - 2; ++i) {
        if (i < n) {
          const T d1 = k[i + p - 1] - k[i];
          const T nn1 = t - k[i];
          const T d2 = k[i + p] - k[i + 1];
          const T nn2 = k[i + p] - t;
          T e1, e2, v;
          if ((double)d1 < 1e-10 && (double)d2 < 1e-10) {
            v = (T)0;
          } else {
            if ((double)d1 < 1e-10) {
              e1 = (T)0;
              e2 = nn2 / d2 * N[i + 1];
            } else if ((double)d2 < 1e-10) {
              e2 = (T)0;
              e1 = nn1 / d1 * N[i];
            } else {
              e1 = nn1 / d1 * N[i];
              e2 = nn2 / d2 * N[i + 1];
            }
            v = (e1 + e2 < (T)0) ? (T)0 : e1 + e2;
          }
          N[i] = v;  // in place: N[i+1] of the previous level is read before it is overwritten
        }
      }
    }
  }
  if (lower && order == 1) {
#pragma unroll
    for (int i = 0; i < NB + 1; ++i) lower[i] = (T)0;
  }
#pragma unroll
  for (int b = 0; b < NB; ++b) out[b] = N[b];
}

template <int NB, int NO>
__device__ __forceinline__ void basis_values(const Spline& sp, float t, float* out, float* lower) {
  basis_values_t<NB, NO, float>(sp.knots, sp.bins, sp.order, t, out, lower);
}

// d N_{b,order} / dt from the order-1 level (exact derivative of the piecewise polynomial).
template <int NB>
__device__ __forceinline__ void basis_derivs(const Spline& sp, const float* lower, float* d) {
  const int p = sp.order;
  const float* k = sp.knots;
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    float v = 0.0f;
    if (b < sp.bins && p > 1) {
      const float d1 = k[b + p - 1] - k[b];
      const float d2 = k[b + p] - k[b + 1];
      if ((double)d1 >= 1e-10) v += lower[b] / d1;
      if ((double)d2 >= 1e-10) v -= lower[b + 1] / d2;
      v = (float)(p - 1) * v;
    }
    d[b] = v;
  }
}

__device__ __forceinline__ uint16_t f32_to_bf16_rne(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf16_to_f32(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

struct Layout {
  int bins, gw, gt;     // bins, genes per group, genes per tile side
  int ks;               // samples per K block of a plane (32; 64 for the int8 digit planes)
  int64_t G, S, Sp;     // genes, samples, padded samples (a multiple of ks)
  int64_t ntiles;       // gene tiles per side
  int64_t rows;         // weight rows allocated = ntiles * tr
  int tr;               // rows per workgroup tile side (160: FP32 kernel, 320: BF16 kernels)
  __host__ __device__ int64_t row_of(int64_t g, int b) const {
    return (g / gw) * kGroupRows + (g % gw) * bins + b;
  }
  // For 10 bins (8 genes per 80-row group) the BF16 planes of the MI kernels with the register
  // epilogue are stored twice, in two row orders inside each group.  Row order R (operand A, the
  // output rows) puts genes 2q and 2q + 1 at the rows 16i + 4q + r (i < 5, r < 4) that MFMA
  // output lanes 16q .. 16q + 15 hold; column order C (operand B) puts bins 5h .. 5h + 4 of gene c
  // at the columns 16j + c + 8h (j < 5) that lane c + 8h holds.  A lane then holds all its
  // values of two gene pairs, half of each pair's table, the other half in lane ^ 8.
  __host__ __device__ int64_t row_of_r(int64_t g, int b) const {
    const int gl = (int)(g % 8), m = 10 * (gl & 1) + b;
    return (g / 8) * kGroupRows + 16 * (m / 4) + 4 * (gl >> 1) + (m & 3);
  }
  __host__ __device__ int64_t row_of_c(int64_t g, int b) const {
    const int gl = (int)(g % 8);
    return (g / 8) * kGroupRows + 16 * (b % 5) + gl + 8 * (b / 5);
  }
  // Element (row, sample) of a K-blocked panel plane: each 80-row group stores its samples in
  // blocks of ks, so one group's K step is one contiguous 80 x ks slice.
  __host__ __device__ int64_t at(int64_t row, int64_t s) const {
    return (((row / kGroupRows) * (Sp / ks) + s / ks) * kGroupRows + row % kGroupRows) * ks + s % ks;
  }
};

// groups_per_tile: 80-row gene groups per workgroup tile side (2 for every MI kernel: 160 rows);
// ks: samples per K block (kKStep, or kKStep8 for the int8 digit planes)
Layout make_layout(int64_t G, int64_t S, int bins, int groups_per_tile = 2, int ks = kKStep) {
  Layout L;
  L.bins = bins;
  L.gw = kGroupRows / bins;
  L.gt = groups_per_tile * L.gw;
  L.tr = groups_per_tile * kGroupRows;
  L.ks = ks;
  L.G = G;
  L.S = S;
  L.Sp = ceil_div(S, ks) * ks;
  L.ntiles = ceil_div(G, L.gt);
  L.rows = L.ntiles * L.tr;
  return L;
}

// ---------------------------------------------------------------- block reductions
template <typename T>
__device__ T block_sum(T v, T* sh) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  __syncthreads();
  if (l == 0) sh[w] = v;
  __syncthreads();
  T r = 0;
  for (int i = 0; i < (int)(blockDim.x >> 6); ++i) r += sh[i];
  return r;
}

// ---------------------------------------------------------------- per-gene preparation
struct PrepOut {
  float* w32;          // group-tiled FP32 weights (nullable)
  uint16_t* whi;       // group-tiled BF16 hi (nullable)
  uint16_t* wlo;       // group-tiled BF16 lo (nullable)
  uint16_t* whi_c;     // non-null: whi / wlo in row order R, these two in row order C (10 bins)
  uint16_t* wlo_c;
  int8_t* q8h;         // int8 digit planes h, l in row order R (non-null: the I8X3 kernel's planes,
  int8_t* q8l;         // 64-sample K blocks) and q8h_c, q8l_c in row order C
  int8_t* q8h_c;
  int8_t* q8l_c;
  float* wref;         // reference layout G x ldw (nullable)
  int64_t ldw;
  double* H64;         // marginal entropy per gene (nullable)
  double* pmarg;       // marginal probabilities G x bins (nullable)
  float* xs;           // scaled rows G x S (nullable)
  float* scl;          // per gene: scale (1/(max-min) as float, or 1) (nullable)
  int* amin;           // first arg-min (nullable)
  int* amax;           // first arg-max (nullable)
  int* flags;          // flags[0] |= 1 on non-finite input
};

// One workgroup per gene: min-max scaling (MinMaxScale, BSplineMI.hpp:81-115), basis weights
// (BSplineWeightsKernel::operator(), :248-262) and the FP64 marginal entropy (entropy1d_2, :204-224).
// TWO: BF16 planes in the row orders R and C (o.whi_c, 10 bins) instead of Layout::row_of.
// I8: the int8 digit planes of the I8X3 kernel (o.q8h, o.q8l; with TWO also o.q8h_c, o.q8l_c in
// the two row orders) instead of BF16 planes.  TWO launches cover every gene slot of the planes
// (L.ntiles * L.gt) and every sample up to Sp: padding slots and samples are written as zeros, so
// those planes need no clearing.
template <int NB, int NO, bool TWO, bool I8>
__global__ __launch_bounds__(kThreads) void bspline_prep_kernel(const float* __restrict__ X, int64_t ldx, Layout L,
                                                                Spline sp, int do_scale, PrepOut o) {
  const int64_t g = blockIdx.x;
  const int tid = threadIdx.x;
  const int64_t S = L.S;
  if constexpr (TWO) {
    if (g >= L.G) {   // a padding slot of the last gene tile
      const int gl = (int)(g % 8);
      const uint32_t grp_base = (uint32_t)((g / 8) * (L.Sp / L.ks));
      for (int64_t s = tid; s < L.Sp; s += kThreads) {
        const uint32_t sbase = (grp_base + (uint32_t)(s / L.ks)) * (uint32_t)(kGroupRows * L.ks) + (uint32_t)(s % L.ks);
        for (int b = 0; b < 10; ++b) {
          const int m = 10 * (gl & 1) + b;
          const uint32_t er = sbase + (uint32_t)(16 * (m >> 2) + 4 * (gl >> 1) + (m & 3)) * L.ks;
          const uint32_t ec = sbase + (uint32_t)(16 * (b % 5) + gl + 8 * (b / 5)) * L.ks;
          if constexpr (I8) {
            o.q8h[er] = 0; o.q8l[er] = 0; o.q8h_c[ec] = 0; o.q8l_c[ec] = 0;
          } else {
            o.whi[er] = 0; o.whi_c[ec] = 0;
            if (o.wlo) { o.wlo[er] = 0; o.wlo_c[ec] = 0; }
          }
        }
      }
      return;
    }
  }
  const float* x = X + g * ldx;
  __shared__ float shf[4];
  __shared__ int shi[4];
  __shared__ double shd[4];
  __shared__ double psum[kMaxBins];

  float mn = S > 0 ? x[0] : 0.0f, mx = mn;
  int bad = 0;
  for (int64_t s = tid; s < S; s += kThreads) {
    const float v = x[s];
    if (!isfinite(v)) bad = 1;
    mn = fminf(mn, v);
    mx = fmaxf(mx, v);
  }
  // block min / max
  for (int off = 32; off > 0; off >>= 1) {
    mn = fminf(mn, __shfl_xor(mn, off));
    mx = fmaxf(mx, __shfl_xor(mx, off));
  }
  __syncthreads();
  if ((tid & 63) == 0) { shf[tid >> 6] = mn; }
  __syncthreads();
  mn = fminf(fminf(shf[0], shf[1]), fminf(shf[2], shf[3]));
  __syncthreads();
  if ((tid & 63) == 0) { shf[tid >> 6] = mx; }
  __syncthreads();
  mx = fmaxf(fmaxf(shf[0], shf[1]), fmaxf(shf[2], shf[3]));
  if (bad) atomicOr(o.flags, 1);

  float scale = 1.0f, offset = 0.0f;
  if (do_scale && mx > mn) {
    scale = (float)(1.0 / (double)(mx - mn));
    offset = mn;
  }
  if (o.scl && tid == 0) o.scl[g] = scale;
  if (o.amin || o.amax) {
    int imn = INT32_MAX, imx = INT32_MAX;
    for (int64_t s = tid; s < S; s += kThreads) {
      if (x[s] == mn && (int)s < imn) imn = (int)s;
      if (x[s] == mx && (int)s < imx) imx = (int)s;
    }
    for (int off = 32; off > 0; off >>= 1) {
      imn = min(imn, __shfl_xor(imn, off));
      imx = min(imx, __shfl_xor(imx, off));
    }
    __syncthreads();
    if ((tid & 63) == 0) shi[tid >> 6] = imn;
    __syncthreads();
    imn = min(min(shi[0], shi[1]), min(shi[2], shi[3]));
    __syncthreads();
    if ((tid & 63) == 0) shi[tid >> 6] = imx;
    __syncthreads();
    imx = min(min(shi[0], shi[1]), min(shi[2], shi[3]));
    if (tid == 0) {
      if (o.amin) o.amin[g] = imn;
      if (o.amax) o.amax[g] = imx;
    }
  }

  const int bins = sp.bins;
  double acc[NB];
#pragma unroll
  for (int b = 0; b < NB; ++b) acc[b] = 0.0;
  float w[NB];
  // TWO: element (row, s) of a plane is sbase(s) + ks * (row within the gene's 80-row group),
  // and that row depends on b and on the gene's place in its group only (Layout::row_of_r,
  // row_of_c); 32-bit offsets, as the planes are kept under 4 GiB.
  const int gl = (int)(g % 8);
  const uint32_t grp_base = (uint32_t)((g / 8) * (L.Sp / L.ks));
  // TWO: samples S .. Sp - 1 get zero weights
  const int64_t s_end = TWO ? L.Sp : S;
  for (int64_t s = tid; s < s_end; s += kThreads) {
    if (TWO && s >= S) {
#pragma unroll
      for (int b = 0; b < NB; ++b) w[b] = 0.0f;
    } else {
      const float t = do_scale ? scale * (x[s] - offset) : x[s];
      if (o.xs) o.xs[g * S + s] = t;
      basis_values<NB, NO>(sp, t, w, nullptr);
    }
    const uint32_t sbase = (grp_base + (uint32_t)(s / L.ks)) * (uint32_t)(kGroupRows * L.ks) + (uint32_t)(s % L.ks);
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      if (b < bins) {
        acc[b] += (double)w[b];
        if constexpr (I8) {
          // w ~= (h + l / 254) / 127: h = rint(127 w) in [0, 127], l = rint(254 (127 w - h))
          const float w127 = w[b] * 127.0f;
          const float hq = rintf(w127);
          const float lq = fminf(fmaxf(rintf((w127 - hq) * 254.0f), -127.0f), 127.0f);
          if constexpr (TWO) {   // two row orders (10 bins)
            const int m = 10 * (gl & 1) + b;
            const uint32_t er = sbase + (uint32_t)(16 * (m >> 2) + 4 * (gl >> 1) + (m & 3)) * kKStep8;
            const uint32_t ec = sbase + (uint32_t)(16 * (b % 5) + gl + 8 * (b / 5)) * kKStep8;
            o.q8h[er] = (int8_t)hq;
            o.q8l[er] = (int8_t)lq;
            o.q8h_c[ec] = (int8_t)hq;
            o.q8l_c[ec] = (int8_t)lq;
          } else {               // one row order (Layout::row_of)
            const int64_t e = L.at(L.row_of(g, b), s);
            o.q8h[e] = (int8_t)hq;
            o.q8l[e] = (int8_t)lq;
          }
        } else if constexpr (TWO) {
          const uint16_t hi = f32_to_bf16_rne(w[b]);
          const int m = 10 * (gl & 1) + b;
          const uint32_t er = sbase + (uint32_t)(16 * (m >> 2) + 4 * (gl >> 1) + (m & 3)) * kKStep;
          const uint32_t ec = sbase + (uint32_t)(16 * (b % 5) + gl + 8 * (b / 5)) * kKStep;
          o.whi[er] = hi;
          o.whi_c[ec] = hi;
          if (o.wlo) {
            const uint16_t lo = f32_to_bf16_rne(w[b] - bf16_to_f32(hi));
            o.wlo[er] = lo;
            o.wlo_c[ec] = lo;
          }
        } else {
          const int64_t e = L.at(L.row_of(g, b), s);
          if (o.w32) o.w32[e] = w[b];
          if (o.whi) {
            const uint16_t hi = f32_to_bf16_rne(w[b]);
            o.whi[e] = hi;
            if (o.wlo) o.wlo[e] = f32_to_bf16_rne(w[b] - bf16_to_f32(hi));
          }
        }
        if (o.wref && s < S) o.wref[g * o.ldw + (int64_t)b * S + s] = w[b];
      }
    }
  }
  // FP64 marginals and entropy
  for (int b = 0; b < bins; ++b) {
    double v = 0.0;
#pragma unroll
    for (int c = 0; c < NB; ++c)
      if (c == b) v = acc[c];
    const double t = block_sum<double>(v, shd);
    if (tid == 0) psum[b] = t;
  }
  __syncthreads();
  if (tid == 0) {
    const double norm = 1.0 / (double)S;
    double H = 0.0;
    for (int b = 0; b < bins; ++b) {
      const double p = psum[b] * norm;
      if (o.pmarg) o.pmarg[g * bins + b] = p;
      if (p > 0.0) H += p * log2(p);
    }
    H = 0.0 - H;
    if (o.H64) o.H64[g] = H;
    if (o.wref) o.wref[g * o.ldw + (int64_t)bins * S] = (float)H;
  }
}

// Reference-layout weight rows -> group-tiled panels (+ FP64 copy of the stored marginal entropy).
// With the int8 digit planes (q8h non-null, row orders R and C) a weight outside [0, 1] sets flag
// bit 1: the digits h = rint(127 w) and l cover that range only.
__global__ void weights_import_kernel(const float* __restrict__ W, int64_t ldw, Layout L, float* w32, uint16_t* whi,
                                      uint16_t* wlo, uint16_t* whi_c, uint16_t* wlo_c, int8_t* q8h, int8_t* q8l,
                                      int8_t* q8h_c, int8_t* q8l_c, double* H64, int* flags) {
  const int64_t g = blockIdx.y;
  const int64_t total = (int64_t)L.bins * L.S;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e <= total; e += (int64_t)gridDim.x * blockDim.x) {
    const float v = W[g * ldw + e];
    if (!isfinite(v)) atomicOr(flags, 1);
    if (e == total) {
      H64[g] = (double)v;
      continue;
    }
    const int b = (int)(e / L.S);
    const int64_t s = e % L.S;
    if (q8h) {
      if (!(v >= 0.0f && v <= 1.0f)) {
        atomicOr(flags, 2);
        continue;
      }
      const float w127 = v * 127.0f;
      const float hq = rintf(w127);
      const float lq = fminf(fmaxf(rintf((w127 - hq) * 254.0f), -127.0f), 127.0f);
      if (q8h_c) {   // two row orders (10 bins)
        const int64_t er = L.at(L.row_of_r(g, b), s), ec = L.at(L.row_of_c(g, b), s);
        q8h[er] = (int8_t)hq;
        q8l[er] = (int8_t)lq;
        q8h_c[ec] = (int8_t)hq;
        q8l_c[ec] = (int8_t)lq;
      } else {
        const int64_t e = L.at(L.row_of(g, b), s);
        q8h[e] = (int8_t)hq;
        q8l[e] = (int8_t)lq;
      }
      continue;
    }
    const int64_t ei = L.at(L.row_of(g, b), s);
    if (w32) w32[ei] = v;
    if (whi) {
      const uint16_t hi = f32_to_bf16_rne(v);
      const uint16_t lo = f32_to_bf16_rne(v - bf16_to_f32(hi));
      if (whi_c) {   // row orders R and C (see Layout::row_of_r)
        const int64_t er = L.at(L.row_of_r(g, b), s), ec = L.at(L.row_of_c(g, b), s);
        whi[er] = hi;
        whi_c[ec] = hi;
        if (wlo) { wlo[er] = lo; wlo_c[ec] = lo; }
      } else {
        whi[ei] = hi;
        if (wlo) wlo[ei] = lo;
      }
    }
  }
}

// ---------------------------------------------------------------- all-pairs MI on MFMA
enum Prec { kFP32 = 0, kBF16 = 1, kBF16X3 = 2 };

struct MIArgs {
  const float* w32;
  const uint16_t* whi;
  const uint16_t* wlo;
  const uint16_t* whi_c;   // operand-B planes in row order C (register-epilogue kernels)
  const uint16_t* wlo_c;
  const int8_t* q8[4];     // int8 digit planes h, l (row order R), h, l (row order C) of the I8X3 kernel
  const double* H64;
  float* MI;
  int64_t ldmi;
  Layout L;
  int64_t ntri;      // number of (ti <= tj) tile pairs
  const uint32_t* order;  // tile coordinates in launch order, (ti << 16) | tj
  float norm;        // (float)1 / (float)S, the reference's FT norm_factor (BSplineMI.hpp:319)
  // output block: MI holds rows r0 .. r0 + nr - 1 and columns c0 .. c0 + nc - 1 of the G x G
  // matrix (0, G, 0, G for the whole matrix)
  int64_t r0, nr, c0, nc;
};

// Stores the MI of gene pair (gi, gj), gi <= gj, at (gi, gj) and (gj, gi) where those fall
// inside the output block.
__device__ __forceinline__ void store_mi(const MIArgs& a, int64_t gi, int64_t gj, float mi) {
  const uint64_t ri = (uint64_t)(gi - a.r0), cj = (uint64_t)(gj - a.c0);
  if (ri < (uint64_t)a.nr && cj < (uint64_t)a.nc) a.MI[(int64_t)ri * a.ldmi + (int64_t)cj] = mi;
  if (gi == gj) return;
  const uint64_t rj = (uint64_t)(gj - a.r0), ci = (uint64_t)(gi - a.c0);
  if (rj < (uint64_t)a.nr && ci < (uint64_t)a.nc) a.MI[(int64_t)rj * a.ldmi + (int64_t)ci] = mi;
}

// LDS layouts.
// * FP32 panels: rows of 32 floats + 4 pad (144 B), staged through registers.
// * BF16 planes: rows of 32 bf16 (64 B, no padding) whose four 16-byte chunks are stored
//   XOR-swizzled, chunk' = chunk ^ (row & 8 ? 3 : 0).  A gfx950 ds_read_b128 serves 16 lanes per
//   pass from 64 four-byte banks; the lanes of one pass are {0-3, 12-15, 20-23, 24-27} and its
//   translates by 4 and by 32 (measured on MI355X).  The MFMA operand read (lane -> row lane & 15,
//   chunk lane >> 4) then covers all 64 banks in every pass.  The planes are filled by
//   global_load_lds_dwordx4, which writes 64 lanes x 16 B contiguously into LDS, so the swizzle is
//   applied to the global source addresses.  The planes of NST K steps sit in a ring of stage
//   buffers; with early release every wave copies its fragments of step k into registers and a
//   barrier hands the buffer back, so the loads of steps k + 1 .. k + NST run under the MFMAs of
//   step k.
constexpr int kLdsF32 = 36;
constexpr int kPanelF32 = kTileRows * kLdsF32;        // floats per FP32 panel
constexpr int kEpiStride = 81;                        // FP32 kernel t-tile row stride (floats)
constexpr int kEpiWave = kGroupRows * kEpiStride;     // floats per wave t-tile
constexpr int kEpiBytes = 2 * kEpiWave * 4;           // two waves' t-tiles at a time

constexpr int kRowB16 = 64;                           // bytes per BF16 plane row (32 samples)
constexpr int kPlaneB16 = kTileRows * kRowB16;        // bytes per 160-row plane K slice (10 KB)
constexpr int kDmaRows = 16;                          // rows covered by one 1 KB direct load
constexpr int kDmaPerPlane = kTileRows / kDmaRows;    // 10
// buffer resource word 3 for raw (untyped, unswizzled) gfx9-family buffer accesses
constexpr int kRsrcWord3 = 0x00020000;
// s_waitcnt immediate (gfx9 encoding) for lgkmcnt(0) with the vector-memory and export counts free
constexpr int kWaitLgkm0 = 0xc07f;

template <int P>
constexpr int mi_planes() { return P == kBF16X3 ? 2 : 1; }
// LDS of the int8 kernel: NST stages of four 10 KB plane slices, reused by the LDS epilogue
constexpr int mi8_smem_bytes(int nst) { return nst * 4 * kPlaneB16 > kEpiBytes ? nst * 4 * kPlaneB16 : kEpiBytes; }
template <int P>
constexpr int mi_stage_bytes() { return 2 * mi_planes<P>() * kPlaneB16; }   // panels A and B
// NST: K steps held in LDS (one being computed, NST - 1 in flight)
template <int P, int NST>
constexpr int mi_smem_bytes() {
  constexpr int body = (P == kFP32) ? 2 * kPanelF32 * 4 : NST * mi_stage_bytes<P>();
  return body > kEpiBytes ? body : kEpiBytes;
}

// Waits until at most n of this wave's vector-memory loads are outstanding (n wave-uniform;
// s_waitcnt takes an immediate, so n is rounded down to one of the counts the MI kernel uses).
__device__ __forceinline__ void wait_vm_at_most(int n) {
  if (n >= 20) asm volatile("s_waitcnt vmcnt(20)" ::: "memory");
  else if (n >= 15) asm volatile("s_waitcnt vmcnt(15)" ::: "memory");
  else if (n >= 10) asm volatile("s_waitcnt vmcnt(10)" ::: "memory");
  else if (n >= 9) asm volatile("s_waitcnt vmcnt(9)" ::: "memory");
  else if (n >= 6) asm volatile("s_waitcnt vmcnt(6)" ::: "memory");
  else if (n >= 5) asm volatile("s_waitcnt vmcnt(5)" ::: "memory");
  else if (n >= 3) asm volatile("s_waitcnt vmcnt(3)" ::: "memory");
  else if (n >= 2) asm volatile("s_waitcnt vmcnt(2)" ::: "memory");
  else if (n >= 1) asm volatile("s_waitcnt vmcnt(1)" ::: "memory");
  else asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
}

// FP32: loads the K-step [k0, k0+32) of one 160-row panel into registers (16-byte chunks, 8 per
// row, 5 per thread) and stores them to the padded LDS panel.
constexpr int kStageF32 = 5;

__device__ __forceinline__ void stage_load_f32(const MIArgs& a, int64_t row0, int64_t k0, v4i (&st)[kStageF32]) {
  const int tid = threadIdx.x;
#pragma unroll
  for (int q = 0; q < kStageF32; ++q) {
    const int idx = tid + kThreads * q;
    const int r = idx >> 3, c = idx & 7;
    st[q] = *reinterpret_cast<const v4i*>(a.w32 + a.L.at(row0 + r, k0) + c * 4);
  }
}

__device__ __forceinline__ void stage_store_f32(char* panel, const v4i (&st)[kStageF32]) {
  const int tid = threadIdx.x;
#pragma unroll
  for (int q = 0; q < kStageF32; ++q) {
    const int idx = tid + kThreads * q;
    const int r = idx >> 3, c = idx & 7;
    *reinterpret_cast<v4i*>(reinterpret_cast<float*>(panel) + r * kLdsF32 + c * 4) = st[q];
  }
}

// BF16: the direct loads of one K step.  Each 1 KB load fills 16 rows x 64 B of one plane slice:
// LDS row 16c + lane / 4, physical chunk lane % 4, from logical chunk (lane % 4) ^ swizzle of the
// K-blocked panel.  A 16-row run never crosses an 80-row group, so the source is one contiguous
// 1 KB.  The loads go through a buffer resource (scalar base and range, hardware bounds check);
// the per-load part of the source offset is scalar, `lbyte` is the lane's part.
typedef __attribute__((address_space(3))) void* lds_ptr;

__device__ __forceinline__ void dma_issue(__amdgpu_buffer_rsrc_t rsrc, uint32_t lbyte, uint32_t sbase, uint32_t grp_stride,
                                          int count, int c0, char* dst) {
  // chunks c0 .. c0 + count - 1 of one plane slice; chunk c covers rows 16c .. 16c + 15, i.e.
  // row block c % 5 of 80-row group c / 5
#pragma unroll
  for (int j = 0; j < kDmaPerPlane; ++j) {
    if (j < count) {
      const int c = c0 + j;
      const uint32_t soff = sbase + (uint32_t)(c / 5) * grp_stride + (uint32_t)(c % 5) * (kDmaRows * kKStep * 2);
      __builtin_amdgcn_raw_ptr_buffer_load_lds(rsrc, (lds_ptr)(dst + j * 1024), 16, lbyte, soff, 0, 0);
    }
  }
}

// Chunks j0 and j0 + 1 of the `count` chunks dma_issue would issue.
__device__ __forceinline__ void dma_issue_part(__amdgpu_buffer_rsrc_t rsrc, uint32_t lbyte, uint32_t sbase,
                                               uint32_t grp_stride, int count, int c0, int j0, char* dst) {
#pragma unroll
  for (int jj = 0; jj < 2; ++jj) {
    const int j = j0 + jj;
    if (j < count) {
      const int c = c0 + j;
      const uint32_t soff = sbase + (uint32_t)(c / 5) * grp_stride + (uint32_t)(c % 5) * (kDmaRows * kKStep * 2);
      __builtin_amdgcn_raw_ptr_buffer_load_lds(rsrc, (lds_ptr)(dst + j * 1024), 16, lbyte, soff, 0, 0);
    }
  }
}

// Register epilogue of the kernels whose planes come in the two row orders (10 bins): lane
// (16q + c + 8h) holds the rows of genes 2q, 2q + 1 (row order R) against bins 5h .. 5h + 4 of
// gene c (row order C); t = p log2 p with p = acc * a.norm is summed per gene pair, the two column
// halves are joined across lanes c and c + 8, and MI = (H_i + H_j) + sum t in FP64.
__device__ __forceinline__ void re_epilogue(const MIArgs& a, const floatx4 (&acc)[5][5], int64_t ti, int64_t tj) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wr = wave >> 1, wc = wave & 1, lr = lane & 15, lq = lane >> 4;
  const bool diag = ti == tj;
  float s0 = 0.0f, s1 = 0.0f;
#pragma unroll
  for (int i = 0; i < 5; ++i)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      float part = 0.0f;
#pragma unroll
      for (int j = 0; j < 5; ++j) {
        const float p = acc[i][j][r] * a.norm;
        part += p > 0.0f ? p * log2f(p) : 0.0f;
      }
      if (4 * i + r < 10) s0 += part;   // row 4i + r of the lane's 20: gene 2q for the first 10
      else s1 += part;
    }
  s0 += __shfl_xor(s0, 8);
  s1 += __shfl_xor(s1, 8);
  if (lr < 8) {
    const int64_t gj = tj * a.L.gt + wc * 8 + lr;
#pragma unroll
    for (int hh = 0; hh < 2; ++hh) {
      const int64_t gi = ti * a.L.gt + wr * 8 + 2 * lq + hh;
      if (gi >= a.L.G || gj >= a.L.G || (diag && gi > gj)) continue;
      const float mi = (float)((a.H64[gi] + a.H64[gj]) + (double)(hh ? s1 : s0));
      store_mi(a, gi, gj, mi);
    }
  }
}

// LDS epilogue of the kernels with one row order (any bin count): each wave's t = p log2 p tile
// goes to LDS (two waves at a time, over the stage buffers), then each lane sums one gene pair's
// bins x bins block, and MI = (H_i + H_j) + sum t in FP64.  Every thread of the block calls it.
__device__ __forceinline__ void lds_epilogue(const MIArgs& a, const floatx4 (&acc)[5][5], int64_t ti, int64_t tj,
                                             char* smem) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wr = wave >> 1, wc = wave & 1, lr = lane & 15, lq = lane >> 4;
  const bool diag = ti == tj;
  float* E = reinterpret_cast<float*>(smem);
  const int bins = a.L.bins, gw = a.L.gw;
  for (int phase = 0; phase < 2; ++phase) {
    if (wr == phase) {
      float* Ew = E + wc * kEpiWave;
#pragma unroll
      for (int i = 0; i < 5; ++i)
#pragma unroll
        for (int j = 0; j < 5; ++j)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const float p = acc[i][j][r] * a.norm;
            const float t = p > 0.0f ? p * log2f(p) : 0.0f;
            Ew[(i * 16 + lq * 4 + r) * kEpiStride + j * 16 + lr] = t;
          }
    }
    __syncthreads();
    if (wr == phase) {
      const float* Ew = E + wc * kEpiWave;
      for (int pr = lane; pr < gw * gw; pr += 64) {
        const int pi = pr / gw, pj = pr % gw;
        const int64_t gi = ti * a.L.gt + wr * gw + pi;
        const int64_t gj = tj * a.L.gt + wc * gw + pj;
        if (gi >= a.L.G || gj >= a.L.G) continue;
        if (diag && gi > gj) continue;
        float sum = 0.0f;
        for (int r = 0; r < bins; ++r) {
          const float* row = Ew + (pi * bins + r) * kEpiStride + pj * bins;
          for (int c = 0; c < bins; ++c) sum += row[c];
        }
        const float mi = (float)((a.H64[gi] + a.H64[gj]) + (double)sum);
        store_mi(a, gi, gj, mi);
      }
    }
    __syncthreads();
  }
}

// All-pairs MI: 160 x 160 workgroup tile (16 x 16 genes for 10 bins), four waves of 80 x 80
// (5 x 5 blocks of 16 x 16 MFMA outputs).  FP32 operands are staged through registers into padded
// LDS panels (two workgroups per CU); BF16 planes arrive through an NST-stage ring of direct loads
// (NST - 1 K steps in flight while one is computed).  RE (10 bins, BF16 planes in row orders R
// and C, see Layout::row_of_r): the epilogue sums each gene pair in registers.
template <int P, int NST, bool ER, bool RE>
__global__ __launch_bounds__(kThreads) void bspline_mi_kernel(MIArgs a) {
  extern __shared__ __align__(16) char smem[];

  int64_t ti, tj;
  {
    const uint32_t code = a.order[blockIdx.x];
    ti = code >> 16;
    tj = code & 0xffffu;
  }
  const bool diag = ti == tj;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = wave >> 1, wc = wave & 1;
  const int lr = lane & 15, lq = lane >> 4;
  const int64_t rowA = ti * kTileRows, rowB = tj * kTileRows;

  floatx4 acc[5][5];
#pragma unroll
  for (int i = 0; i < 5; ++i)
#pragma unroll
    for (int j = 0; j < 5; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};

  const int64_t nk = a.L.Sp / kKStep;
  if constexpr (P == kFP32) {
    char* panelA = smem;
    char* panelB = smem + kPanelF32 * 4;
    v4i sa[kStageF32], sb[kStageF32];
    stage_load_f32(a, rowA, 0, sa);
    if (!diag) stage_load_f32(a, rowB, 0, sb);
    for (int64_t kk = 0; kk < nk; ++kk) {
      stage_store_f32(panelA, sa);
      if (!diag) stage_store_f32(panelB, sb);
      __syncthreads();
      if (kk + 1 < nk) {
        stage_load_f32(a, rowA, (kk + 1) * kKStep, sa);
        if (!diag) stage_load_f32(a, rowB, (kk + 1) * kKStep, sb);
      }
      const float* As = reinterpret_cast<const float*>(panelA);
      const float* Bs = reinterpret_cast<const float*>(diag ? panelA : panelB);
      // B fragments stay in registers for the whole K step; A fragments are read per block row.
      floatx4 bfr[5][2];
  #pragma unroll
      for (int j = 0; j < 5; ++j) {
        const float* pb = Bs + (wc * 80 + j * 16 + lr) * kLdsF32 + lq * 8;
        bfr[j][0] = *reinterpret_cast<const floatx4*>(pb);
        bfr[j][1] = *reinterpret_cast<const floatx4*>(pb + 4);
      }
  #pragma unroll
      for (int i = 0; i < 5; ++i) {
        const float* pa = As + (wr * 80 + i * 16 + lr) * kLdsF32 + lq * 8;
        const floatx4 a0 = *reinterpret_cast<const floatx4*>(pa);
        const floatx4 a1 = *reinterpret_cast<const floatx4*>(pa + 4);
  #pragma unroll
        for (int u = 0; u < 8; ++u)
  #pragma unroll
          for (int j = 0; j < 5; ++j)
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(u < 4 ? a0[u & 3] : a1[u & 3], bfr[j][u >> 2][u & 3],
                                                              acc[i][j], 0, 0, 0);
      }
      __syncthreads();
    }
  } else {
    constexpr int np = mi_planes<P>();
    constexpr int kStage = mi_stage_bytes<P>();
    // Load assignment: the stage's plane slices (A.hi, [A.lo,] B.hi, [B.lo]; A only on diagonal
    // tiles) are split evenly over the four waves; each wave takes a contiguous run of one
    // slice's 16-row chunks.
    const int swave = __builtin_amdgcn_readfirstlane(wave);
    // a diagonal tile reads one panel for both operands, unless they come in two row orders
    const bool one_panel = diag && !RE;
    const int nslices = (one_panel ? 1 : 2) * np;              // 1, 2 or 4
    const int wps = 4 / nslices;                               // waves per slice
    const int slice = swave / wps;
    const int nper = (kDmaPerPlane + wps - 1) / wps;
    const int c0 = (swave % wps) * nper;
    const int count = kDmaPerPlane - c0 < nper ? kDmaPerPlane - c0 : nper;
    const int lrow = lane >> 2;
    const uint32_t lbyte = (lrow * kKStep + ((((lane & 3) ^ ((lrow & 8) ? 3 : 0))) << 3)) * 2;
    const uint32_t plane_bytes = (uint32_t)(a.L.rows * a.L.Sp * 2);
    const bool lo_plane = P == kBF16X3 && (slice % np) == 1;
    const uint16_t* plane = (RE && slice / np == 1) ? (lo_plane ? a.wlo_c : a.whi_c) : (lo_plane ? a.wlo : a.whi);
    const __amdgpu_buffer_rsrc_t rsrc =
        __builtin_amdgcn_make_buffer_rsrc((void*)plane, (short)0, plane_bytes, kRsrcWord3);
    const uint32_t grp_stride = (uint32_t)(nk * kGroupRows * kKStep * 2);      // bytes between groups
    const uint32_t kstep_bytes = (uint32_t)(kGroupRows * kKStep * 2);
    const uint32_t sbase0 = (uint32_t)(((slice / np) == 0 ? rowA : rowB) / kGroupRows) * grp_stride;
    const int dst_off = slice * kPlaneB16 + c0 * 1024;
    // byte offset of this lane's MFMA fragment within a 16-row block of a swizzled plane
    const int frag = lr * kRowB16 + ((lq ^ ((lr & 8) ? 3 : 0)) << 4);

    // A ring of NST stage buffers.  Without early release (ER) a buffer is read fragment by
    // fragment during its step's MFMAs, so NST - 1 steps are in flight.  With ER every wave copies
    // all of its step's fragments into registers first and a second barrier frees the buffer for
    // the loads of step kk + NST, so all NST buffers are in flight during the MFMAs.
    constexpr int kAhead = ER ? NST : NST - 1;   // steps issued before step 0 is computed
#pragma unroll
    for (int s = 0; s < kAhead; ++s)
      if (s < nk) dma_issue(rsrc, lbyte, sbase0 + (uint32_t)s * kstep_bytes, grp_stride, count, c0, smem + s * kStage + dst_off);
    for (int64_t kk = 0; kk < nk; ++kk) {
      // this wave's loads of step kk have landed once only the loads of the later steps already
      // issued (at most kAhead - 1) are outstanding; after the barrier everyone's have
      const int64_t later = nk - 1 - kk < kAhead - 1 ? nk - 1 - kk : kAhead - 1;
      wait_vm_at_most((int)later * count);
      // a bare barrier: __syncthreads()'s release fence would wait for every load in flight
      asm volatile("s_barrier" ::: "memory");
      const char* st = smem + (kk % NST) * kStage;
      // The loads of a later step are issued a few chunks at a time between the MFMA block rows
      // below, so that a wave never waits long on a full load queue with MFMA work at hand.
      // Without ER they fill the buffer of step kk - 1, which every wave finished before the
      // barrier; with ER the buffer of step kk, once the second barrier has freed it.
      const int64_t next = ER ? kk + NST : kk + NST - 1;
      const uint32_t next_src = sbase0 + (uint32_t)next * kstep_bytes;
      char* const next_dst = smem + (ER ? kk % NST : (kk + NST - 1) % NST) * kStage + dst_off;
      const int next_count = next < nk ? count : 0;
      const char* sA = st;
      const char* sB = one_panel ? st : st + np * kPlaneB16;
      bf16x8 bh[5], bl[5], ahs[5], als[5];
      // the three products (two for the low halves, then hi x hi) of block row i
      auto mfma_row = [&](int i, const bf16x8& ah, const bf16x8& al) {
        if (P == kBF16X3) {
#pragma unroll
          for (int j = 0; j < 5; ++j) {
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh[j], acc[i][j], 0, 0, 0);
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl[j], acc[i][j], 0, 0, 0);
          }
        }
#pragma unroll
        for (int j = 0; j < 5; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh[j], acc[i][j], 0, 0, 0);
      };
      if (ER) {
        // Every fragment of the step is copied into registers, those of block row 0 first (B hi,
        // A lo 0, B lo, A hi 0, then rows 1-4), each read held in place by a scheduling barrier,
        // so that the MFMAs of rows 0 and 1 start as their operands arrive.  The barrier that
        // frees the buffer follows them, and the next step's loads go out over rows 2-4.
        // Nothing is outstanding in the LDS/scalar counter here; the explicit wait says so to the
        // wait-count pass, which otherwise (a scalar load of the epilogue's arguments still in
        // flight from the prologue) waits for all twenty reads before the first MFMA.
        __builtin_amdgcn_s_waitcnt(kWaitLgkm0);
        const int oa = (wr * 80) * kRowB16 + frag, ob = (wc * 80) * kRowB16 + frag;
        bh[0] = *reinterpret_cast<const bf16x8*>(sB + ob);
        if (P == kBF16X3) als[0] = *reinterpret_cast<const bf16x8*>(sA + kPlaneB16 + oa);
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int j = 1; j < 5; ++j) {
          bh[j] = *reinterpret_cast<const bf16x8*>(sB + ob + j * 16 * kRowB16);
          __builtin_amdgcn_sched_barrier(0);
        }
        if (P == kBF16X3) {
#pragma unroll
          for (int j = 0; j < 5; ++j) {
            bl[j] = *reinterpret_cast<const bf16x8*>(sB + kPlaneB16 + ob + j * 16 * kRowB16);
            __builtin_amdgcn_sched_barrier(0);
          }
        }
        ahs[0] = *reinterpret_cast<const bf16x8*>(sA + oa);
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int i = 1; i < 5; ++i) {
          const int o = oa + i * 16 * kRowB16;
          if (P == kBF16X3) als[i] = *reinterpret_cast<const bf16x8*>(sA + kPlaneB16 + o);
          ahs[i] = *reinterpret_cast<const bf16x8*>(sA + o);
        }
        __builtin_amdgcn_sched_barrier(0);
        constexpr int kRelease = 2;                                            // block row
        constexpr int kPer = (kDmaPerPlane + 4 - kRelease) / (5 - kRelease);   // chunks per later row
#pragma unroll
        for (int i = 0; i < 5; ++i) {
          if (i == kRelease) {
            // every wave's reads of this buffer are complete after the barrier
            __builtin_amdgcn_sched_barrier(0);
            asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
            __builtin_amdgcn_sched_barrier(0);
          }
          if (i >= kRelease) {
#pragma unroll
            for (int jj = 0; jj < kPer; ++jj) {
              const int j = (i - kRelease) * kPer + jj;
              if (j < next_count) {
                const int c = c0 + j;
                const uint32_t soff = next_src + (uint32_t)(c / 5) * grp_stride + (uint32_t)(c % 5) * (kDmaRows * kKStep * 2);
                __builtin_amdgcn_raw_ptr_buffer_load_lds(rsrc, (lds_ptr)(next_dst + j * 1024), 16, lbyte, soff, 0, 0);
              }
            }
          }
          mfma_row(i, ahs[i], als[i]);
        }
      } else {
        // B fragments stay in registers for the whole K step; A fragments are read per block row
#pragma unroll
        for (int j = 0; j < 5; ++j) {
          const int o = (wc * 80 + j * 16) * kRowB16 + frag;
          bh[j] = *reinterpret_cast<const bf16x8*>(sB + o);
          if (P == kBF16X3) bl[j] = *reinterpret_cast<const bf16x8*>(sB + kPlaneB16 + o);
        }
#pragma unroll
        for (int i = 0; i < 5; ++i) {
          dma_issue_part(rsrc, lbyte, next_src, grp_stride, next_count, c0, 2 * i, next_dst);
          const int o = (wr * 80 + i * 16) * kRowB16 + frag;
          const bf16x8 ah = *reinterpret_cast<const bf16x8*>(sA + o);
          const bf16x8 al = P == kBF16X3 ? *reinterpret_cast<const bf16x8*>(sA + kPlaneB16 + o) : ah;
          mfma_row(i, ah, al);
        }
      }
    }
    __syncthreads();
  }

  if constexpr (RE) {
    re_epilogue(a, acc, ti, tj);
    return;
  }
  lds_epilogue(a, acc, ti, tj, smem);
}

// I8X3: each weight is split into two int8 digits, w ~= (h + l / 254) / 127 with
// h = rint(127 w) in [0, 127] and l = rint(254 (127 w - h)) in [-127, 127] (absolute error of the
// split <= 1 / 64516 + 1 ulp).  Per 64-sample K step, v_mfma_i32_16x16x64_i8 accumulates h_i h_j
// into one int32 set and h_i l_j + l_i h_j into a second; both sums are exact (|HH| <= 127^2 Sp,
// |X| <= 2 * 127^2 Sp, below 2^31 for Sp <= 66560).  The epilogue forms p = (HH + X / 254) / (127^2 S);
// the dropped l_i l_j term is below 1 / 254^2 of a digit product.  The int8 MFMA runs at twice the
// BF16 rate and each digit plane takes one byte per sample, so a K step of 64 samples moves the
// bytes and costs the MFMA cycles of a 32-sample BF16 step with half of BF16X3's products.
// Same tile (160 x 160, four 80 x 80 waves), direct loads, swizzled 64-byte LDS rows and
// register epilogue as the BF16 kernels with two row orders; the two accumulator sets (200 of
// the 256 registers a wave has at two waves per SIMD) leave no room for a step's A fragments, so
// a K step's buffer stays in LDS until its last block row and NST - 1 steps are in flight.
// EARLY: the next step's ten loads are issued together after the barrier (else two per block row).
// RE: planes in the two row orders and the register epilogue (10 bins); otherwise one row order
// (a.q8[2], a.q8[3] repeat a.q8[0], a.q8[1]) and the LDS epilogue.
template <int NST, bool EARLY, bool RE>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(NST == 2 ? 2 : 1, 2))) void bspline_mi_i8_kernel(MIArgs a) {
  extern __shared__ __align__(16) char smem[];

  int64_t ti, tj;
  {
    const uint32_t code = a.order[blockIdx.x];
    ti = code >> 16;
    tj = code & 0xffffu;
  }
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = wave >> 1, wc = wave & 1;
  const int lr = lane & 15, lq = lane >> 4;
  const int64_t rowA = ti * kTileRows, rowB = tj * kTileRows;

  v4i acc_h[5][5], acc_x[5][5];
#pragma unroll
  for (int i = 0; i < 5; ++i)
#pragma unroll
    for (int j = 0; j < 5; ++j) {
      acc_h[i][j] = v4i{0, 0, 0, 0};
      acc_x[i][j] = v4i{0, 0, 0, 0};
    }

  const int64_t nk = a.L.Sp / kKStep8;
  constexpr int kStage = 4 * kPlaneB16;    // A.h, A.l (order R), B.h, B.l (order C): 40 KB
  // load assignment: wave w fills plane slice w of each stage (ten 1 KB direct loads)
  const int slice = __builtin_amdgcn_readfirstlane(wave);
  constexpr int count = kDmaPerPlane;
  const int lrow = lane >> 2;
  const uint32_t lbyte = (uint32_t)(lrow * 64 + (((lane & 3) ^ ((lrow & 8) ? 3 : 0)) << 4));
  const uint32_t plane_bytes = (uint32_t)(a.L.rows * a.L.Sp);
  const __amdgpu_buffer_rsrc_t rsrc =
      __builtin_amdgcn_make_buffer_rsrc((void*)a.q8[slice], (short)0, plane_bytes, kRsrcWord3);
  const uint32_t grp_stride = (uint32_t)(nk * kGroupRows * 64);     // bytes between 80-row groups
  constexpr uint32_t kstep_bytes = kGroupRows * 64;
  const uint32_t sbase0 = (uint32_t)((slice < 2 ? rowA : rowB) / kGroupRows) * grp_stride;
  const int dst_off = slice * kPlaneB16;
  const int frag = lr * kRowB16 + ((lq ^ ((lr & 8) ? 3 : 0)) << 4);
  const int oa = (wr * 80) * kRowB16 + frag, ob = (wc * 80) * kRowB16 + frag;

  constexpr int kAhead = NST - 1;
#pragma unroll
  for (int st = 0; st < kAhead; ++st)
    if (st < nk) dma_issue(rsrc, lbyte, sbase0 + (uint32_t)st * kstep_bytes, grp_stride, count, 0, smem + st * kStage + dst_off);
  for (int64_t kk = 0; kk < nk; ++kk) {
    const int64_t later = nk - 1 - kk < kAhead - 1 ? nk - 1 - kk : kAhead - 1;
    wait_vm_at_most((int)later * count);
    asm volatile("s_barrier" ::: "memory");
    const char* sA = smem + (kk % NST) * kStage;
    const char* sB = sA + 2 * kPlaneB16;
    // the buffer of step kk - 1, which every wave finished before the barrier, takes step kk + NST - 1
    const int64_t next = kk + NST - 1;
    const uint32_t next_src = sbase0 + (uint32_t)next * kstep_bytes;
    char* const next_dst = smem + (next % NST) * kStage + dst_off;
    const int next_count = next < nk ? count : 0;
    if (EARLY) dma_issue(rsrc, lbyte, next_src, grp_stride, next_count, 0, next_dst);
    v4i bh[5], bl[5];
#pragma unroll
    for (int j = 0; j < 5; ++j) {
      bh[j] = *reinterpret_cast<const v4i*>(sB + ob + j * 16 * kRowB16);
      bl[j] = *reinterpret_cast<const v4i*>(sB + kPlaneB16 + ob + j * 16 * kRowB16);
    }
#pragma unroll
    for (int i = 0; i < 5; ++i) {
      if (!EARLY) dma_issue_part(rsrc, lbyte, next_src, grp_stride, next_count, 0, 2 * i, next_dst);
      const v4i ah = *reinterpret_cast<const v4i*>(sA + oa + i * 16 * kRowB16);
      const v4i al = *reinterpret_cast<const v4i*>(sA + kPlaneB16 + oa + i * 16 * kRowB16);
#pragma unroll
      for (int j = 0; j < 5; ++j) acc_h[i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(ah, bh[j], acc_h[i][j], 0, 0, 0);
#pragma unroll
      for (int j = 0; j < 5; ++j) acc_x[i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(ah, bl[j], acc_x[i][j], 0, 0, 0);
#pragma unroll
      for (int j = 0; j < 5; ++j) acc_x[i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(al, bh[j], acc_x[i][j], 0, 0, 0);
    }
  }
  floatx4 acc[5][5];
#pragma unroll
  for (int i = 0; i < 5; ++i)
#pragma unroll
    for (int j = 0; j < 5; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) acc[i][j][r] = (float)acc_h[i][j][r] + (float)acc_x[i][j][r] * (1.0f / 254.0f);
  if constexpr (RE) {
    re_epilogue(a, acc, ti, tj);
  } else {
    __syncthreads();   // every wave is done with the stage buffers, which the epilogue reuses
    lds_epilogue(a, acc, ti, tj, smem);
  }
}

// ---------------------------------------------------------------- FP64 (the reference's double DataType)
// MinMaxScale<double>, BSplineWeightsKernel<double> and BSplineMIKernel<double>
// (BSplineMI.hpp:75-116, :119-274, :339-500): scaling, basis, marginals and joint tables in
// double.  The joint tables W_i W_j^T run on v_mfma_f64_16x16x4_f64.
constexpr int kKStep64 = 8;   // samples per K block of the FP64 weight panel (64-byte rows)
typedef double doublex2 __attribute__((ext_vector_type(2)));
typedef double doublex4 __attribute__((ext_vector_type(4)));

// buildKnotVector<double> (BSplineMI.hpp:35-71)
struct Spline64 {
  int bins, order;
  double knots[kMaxBins + kMaxOrder];
};

Spline64 make_spline64(int bins, int order) {
  Spline64 sp;
  std::memset(&sp, 0, sizeof sp);
  sp.bins = bins;
  sp.order = order;
  const int n_int = bins - order;
  const double norm = 1.0 / (double)(n_int + 1);
  for (int i = 0; i < bins + order; ++i) {
    if (i < order) sp.knots[i] = 0.0;
    else if (i < order + n_int) sp.knots[i] = (double)(i - order + 1) * norm;
    else sp.knots[i] = 1.0;
  }
  return sp;
}

struct Prep64Out {
  double* w;        // group-tiled weights, Layout with ks = kKStep64 (nullable)
  double* wref;     // reference layout: G rows of bins*S weights (bin-major) + H (nullable)
  int64_t ldw;
  double* xs;       // scaled rows (nullable)
  int64_t ldxs;
  double* H64;      // marginal entropy per gene (nullable)
  int* flags;       // flags[0] |= 1 on non-finite input
};

// One workgroup per gene: MinMaxScale<double> (scale = 1.0 / (max - min), x' = scale (x - min);
// a constant row keeps its values), the basis in double and entropy1d_2's marginal entropy
// (ex += w * (1.0 / S), H = 0 - sum ex log2 ex).
template <int NB, int NO>
__global__ __launch_bounds__(kThreads) void bspline_prep64_kernel(const double* __restrict__ X, int64_t ldx, Layout L,
                                                                  Spline64 sp, int do_scale, Prep64Out o) {
  const int64_t g = blockIdx.x;
  const int tid = threadIdx.x;
  const int64_t S = L.S;
  const double* x = X + g * ldx;
  __shared__ double shd[4];
  __shared__ double psum[kMaxBins];

  double mn = S > 0 ? x[0] : 0.0, mx = mn;
  int bad = 0;
  for (int64_t s = tid; s < S; s += kThreads) {
    const double v = x[s];
    if (!isfinite(v)) bad = 1;
    mn = fmin(mn, v);
    mx = fmax(mx, v);
  }
  for (int off = 32; off > 0; off >>= 1) {
    mn = fmin(mn, __shfl_xor(mn, off));
    mx = fmax(mx, __shfl_xor(mx, off));
  }
  __syncthreads();
  if ((tid & 63) == 0) shd[tid >> 6] = mn;
  __syncthreads();
  mn = fmin(fmin(shd[0], shd[1]), fmin(shd[2], shd[3]));
  __syncthreads();
  if ((tid & 63) == 0) shd[tid >> 6] = mx;
  __syncthreads();
  mx = fmax(fmax(shd[0], shd[1]), fmax(shd[2], shd[3]));
  if (bad) atomicOr(o.flags, 1);

  double scale = 1.0, offset = 0.0;
  if (do_scale && mx > mn) {
    scale = 1.0 / (mx - mn);
    offset = mn;
  }
  const int bins = sp.bins;
  const double norm = 1.0 / (double)S;
  double acc[NB];
#pragma unroll
  for (int b = 0; b < NB; ++b) acc[b] = 0.0;
  double w[NB];
  for (int64_t s = tid; s < S; s += kThreads) {
    const double t = do_scale ? scale * (x[s] - offset) : x[s];
    if (o.xs) o.xs[g * o.ldxs + s] = t;
    basis_values_t<NB, NO, double>(sp.knots, bins, sp.order, t, w, nullptr);
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      if (b < bins) {
        acc[b] += w[b] * norm;
        if (o.w) o.w[L.at(L.row_of(g, b), s)] = w[b];
        if (o.wref) o.wref[g * o.ldw + (int64_t)b * S + s] = w[b];
      }
    }
  }
  for (int b = 0; b < bins; ++b) {
    double v = 0.0;
#pragma unroll
    for (int c = 0; c < NB; ++c)
      if (c == b) v = acc[c];
    const double t = block_sum<double>(v, shd);
    if (tid == 0) psum[b] = t;
  }
  __syncthreads();
  if (tid == 0) {
    double H = 0.0;
    for (int b = 0; b < bins; ++b)
      if (psum[b] > 0.0) H += psum[b] * log2(psum[b]);
    H = 0.0 - H;
    if (o.H64) o.H64[g] = H;
    if (o.wref) o.wref[g * o.ldw + (int64_t)bins * S] = H;
  }
}

// Reference-layout FP64 weight rows (bins*S weights, then H) -> group-tiled panel + H64.
__global__ void weights_import64_kernel(const double* __restrict__ W, int64_t ldw, Layout L, double* __restrict__ w,
                                        double* __restrict__ H64, int* __restrict__ flags) {
  const int64_t n = (int64_t)L.bins * L.S;
  bool bad = false;
  for (int64_t g = blockIdx.y; g < L.G; g += gridDim.y) {
    const double* row = W + g * ldw;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e <= n; e += (int64_t)gridDim.x * blockDim.x) {
      const double v = row[e];
      bad |= !isfinite(v);
      if (e == n) H64[g] = v;
      else w[L.at(L.row_of(g, (int)(e / L.S)), e % L.S)] = v;
    }
  }
  if (__ballot(bad) != 0 && (threadIdx.x & 63) == 0) atomicOr(flags, 1);
}

struct MI64Args {
  const double* w;        // group-tiled weights (Layout with ks = kKStep64, zero padding)
  const double* H64;
  double* MI;
  int64_t ldmi;
  Layout L;
  const uint32_t* order;  // tile coordinates in launch order, (ti << 16) | tj
  double norm;            // 1.0 / S: BSplineEntropyKernel<double>::norm_factor (BSplineMI.hpp:319)
  // output block: MI holds rows r0 .. r0 + nr - 1 and columns c0 .. c0 + nc - 1 of the G x G
  // matrix (0, G, 0, G for the whole matrix)
  int64_t r0, nr, c0, nc;
};

constexpr int kEpi64Stride = 81;                          // doubles per row of a wave's t tile
constexpr int kEpi64Wave = kGroupRows * kEpi64Stride;     // doubles per wave t tile
constexpr int kEpi64Bytes = 2 * kEpi64Wave * 8;           // two waves' t tiles at a time (101.25 KB)

// one 8-sample K block of a wave's 80 x 80 block of MFMA outputs: the .x halves of the fragments
// carry samples 2q, the .y halves samples 2q + 1 (q = lane >> 4, the MFMA's K index)
__device__ __forceinline__ void mi64_step(const doublex2 (&fa)[5], const doublex2 (&fb)[5], doublex4 (&acc)[5][5]) {
#pragma unroll
  for (int i = 0; i < 5; ++i)
#pragma unroll
    for (int j = 0; j < 5; ++j) acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[i].x, fb[j].x, acc[i][j], 0, 0, 0);
#pragma unroll
  for (int i = 0; i < 5; ++i)
#pragma unroll
    for (int j = 0; j < 5; ++j) acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(fa[i].y, fb[j].y, acc[i][j], 0, 0, 0);
}

// fragments of K block kb: block row i of the wave's A (B) group, lane (lr, lq) -> row 16 i + lr,
// samples 2 lq and 2 lq + 1 (one 16-byte load; 64 lanes read one contiguous 1 KB row block)
__device__ __forceinline__ void mi64_load(const double* pa, const double* pb, int64_t kb, doublex2 (&fa)[5],
                                          doublex2 (&fb)[5]) {
  const int64_t off = kb * (kGroupRows * kKStep64);
#pragma unroll
  for (int i = 0; i < 5; ++i) {
    fa[i] = *reinterpret_cast<const doublex2*>(pa + off + i * 16 * kKStep64);
    fb[i] = *reinterpret_cast<const doublex2*>(pb + off + i * 16 * kKStep64);
  }
}

// t = p log2 p in double with p = dot * (1 / S) (BSplineEntropyKernel::entropy2d_simd's
// hx *= norm_factor), summed per gene pair over bins x bins in the reference's (x bin, y bin)
// order through LDS (two waves at a time), and MI = (H_i + H_j) - (0 - sum t)
// (BSplineMIKernel::operator(), BSplineMI.hpp:492-499).  Every thread of the block calls it.
__device__ __forceinline__ void mi64_epilogue(const MI64Args& a, const doublex4 (&acc)[5][5], int64_t ti, int64_t tj,
                                              char* smem) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wr = wave >> 1, wc = wave & 1, lr = lane & 15, lq = lane >> 4;
  double* E = reinterpret_cast<double*>(smem);
  const bool diag = ti == tj;
  const int bins = a.L.bins, gw = a.L.gw;
  for (int phase = 0; phase < 2; ++phase) {
    if (wr == phase) {
      double* Ew = E + wc * kEpi64Wave;
#pragma unroll
      for (int i = 0; i < 5; ++i)
#pragma unroll
        for (int j = 0; j < 5; ++j)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            // v_mfma_f64_16x16x4_f64 output: register r of lane (lr, lq) is D[4 r + lq][lr]
            // (the FP32 16 x 16 forms hold D[4 lq + r][lr]; measured on MI355X)
            const double p = acc[i][j][r] * a.norm;
            Ew[(i * 16 + r * 4 + lq) * kEpi64Stride + j * 16 + lr] = p > 0.0 ? p * log2(p) : 0.0;
          }
    }
    __syncthreads();
    if (wr == phase) {
      const double* Ew = E + wc * kEpi64Wave;
      for (int pr = lane; pr < gw * gw; pr += 64) {
        const int pi = pr / gw, pj = pr % gw;
        const int64_t gi = ti * a.L.gt + wr * gw + pi;
        const int64_t gj = tj * a.L.gt + wc * gw + pj;
        if (gi >= a.L.G || gj >= a.L.G || (diag && gi > gj)) continue;
        double sum = 0.0;
        for (int r = 0; r < bins; ++r) {
          const double* row = Ew + (pi * bins + r) * kEpi64Stride + pj * bins;
          for (int c = 0; c < bins; ++c) sum += row[c];
        }
        const double mi = (a.H64[gi] + a.H64[gj]) - (0.0 - sum);
        // (gi, gj) and (gj, gi) where they fall inside the output block
        const uint64_t ri = (uint64_t)(gi - a.r0), cj = (uint64_t)(gj - a.c0);
        if (ri < (uint64_t)a.nr && cj < (uint64_t)a.nc) a.MI[(int64_t)ri * a.ldmi + (int64_t)cj] = mi;
        const uint64_t rj = (uint64_t)(gj - a.r0), ci = (uint64_t)(gi - a.c0);
        if (gi != gj && rj < (uint64_t)a.nr && ci < (uint64_t)a.nc) a.MI[(int64_t)rj * a.ldmi + (int64_t)ci] = mi;
      }
    }
    __syncthreads();
  }
}

// All-pairs FP64 MI: the 160 x 160 workgroup tile of the other MI kernels, four waves of 80 x 80
// (5 x 5 blocks of 16 x 16 FP64 MFMA outputs, 200 registers).  The FP64 MFMA takes 64 cycles
// per 16 x 16 x 4 product, so operands come straight from memory into a ring of NBUF register
// sets (NBUF K blocks in flight, one wave per SIMD).  Staging them through an LDS ring of direct
// loads instead, with a barrier per K block, measured 1.5-1.6x slower at one wave per SIMD.
// Epilogue: mi64_epilogue.
template <int NBUF>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(1, 1))) void bspline_mi64_kernel(MI64Args a) {
  extern __shared__ __align__(16) char smem[];
  const uint32_t code = a.order[blockIdx.x];
  const int64_t ti = code >> 16, tj = code & 0xffffu;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = wave >> 1, wc = wave & 1, lr = lane & 15, lq = lane >> 4;
  const int64_t nk = a.L.Sp / kKStep64;
  const int64_t group_elems = nk * (kGroupRows * kKStep64);
  const double* pa = a.w + (ti * 2 + wr) * group_elems + lr * kKStep64 + 2 * lq;
  const double* pb = a.w + (tj * 2 + wc) * group_elems + lr * kKStep64 + 2 * lq;

  doublex4 acc[5][5];
#pragma unroll
  for (int i = 0; i < 5; ++i)
#pragma unroll
    for (int j = 0; j < 5; ++j) acc[i][j] = doublex4{0.0, 0.0, 0.0, 0.0};
  // a ring of NBUF register sets: K block kb + NBUF loads into the set K block kb just used
  doublex2 fa[NBUF][5], fb[NBUF][5];
#pragma unroll
  for (int u = 0; u < NBUF; ++u)
    if (u < nk) mi64_load(pa, pb, u, fa[u], fb[u]);
  for (int64_t kb = 0; kb < nk; kb += NBUF) {
#pragma unroll
    for (int u = 0; u < NBUF; ++u) {
      if (kb + u < nk) {
        mi64_step(fa[u], fb[u], acc);
        if (kb + u + NBUF < nk) mi64_load(pa, pb, kb + u + NBUF, fa[u], fb[u]);
      }
    }
  }

  mi64_epilogue(a, acc, ti, tj, smem);
}

// ---------------------------------------------------------------- adjoint
// Per masked pair: joint table p = W_i W_j^T / S (FP32, one wave per pair), its entropy term and
// Gm[b][c] = m * phi(p_bc), phi(p) = -(log2 p + 1/ln 2) = dH/dp for p > 0.
struct PairArgs {
  const float* w32;
  Layout L;
  const int* pi;
  const int* pj;
  const float* m;
  int64_t npairs;
  const double* H64;
  float norm;
  float* Gm;        // npairs x bins*bins
  double* Lpart;    // per pair m * MI
};

template <int NB>
__global__ __launch_bounds__(256) void bspline_pair_adjoint_kernel(PairArgs a) {
  const int lane = threadIdx.x & 63;
  const int64_t n = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (n >= a.npairs) return;
  const int64_t i = a.pi[n], j = a.pj[n];
  const int bins = a.L.bins;
  // each lane accumulates the full table over its samples, then a wave reduction per entry
  float part[NB];  // reused per row b
  double Lsum = 0.0;
  for (int b = 0; b < bins; ++b) {
    const int64_t ri = a.L.row_of(i, b);
#pragma unroll
    for (int c = 0; c < NB; ++c) part[c] = 0.0f;
    for (int64_t s = lane; s < a.L.S; s += 64) {
      const float x = a.w32[a.L.at(ri, s)];
      if (x == 0.0f) continue;
#pragma unroll
      for (int c = 0; c < NB; ++c)
        if (c < bins) part[c] += x * a.w32[a.L.at(a.L.row_of(j, c), s)];
    }
#pragma unroll
    for (int c = 0; c < NB; ++c) {
      if (c >= bins) continue;
      float v = part[c];
      for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
      const float p = v * a.norm;
      float gm = 0.0f;
      if (p > 0.0f) {
        Lsum += (double)(p * log2f(p));
        gm = a.m[n] * (-(log2f(p) + 1.4426950408889634f));
      }
      if (lane == 0) a.Gm[n * bins * bins + b * bins + c] = gm;
    }
  }
  if (lane == 0) a.Lpart[n] = (double)a.m[n] * ((a.H64[i] + a.H64[j]) + Lsum);
}

// The same, for bins <= 16, on MFMA: the pair's table W_i W_j^T is one 16 x 16 output of
// v_mfma_f32_16x16x4_f32 over the samples (rows: gene i's bins, columns: gene j's, zero
// beyond bins).  Lane (r, k) supplies W_i[r][s] and W_j[r][s] for s = s0 + 4k + t in the t-th
// product of a 16-sample step, from one float4 of each panel row; output register q of lane
// (r, k) is table entry (4k + q, r).
__global__ __launch_bounds__(256) void bspline_pair_adjoint_mfma_kernel(PairArgs a) {
  const int lane = threadIdx.x & 63;
  const int64_t n = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (n >= a.npairs) return;
  const int64_t i = a.pi[n], j = a.pj[n];
  const int bins = a.L.bins;
  const int lr = lane & 15, lk = lane >> 4;
  const bool rowok = lr < bins;
  const int64_t step = (int64_t)kGroupRows * kKStep;   // floats per 32-sample block of a group
  const int64_t nkb = a.L.Sp / kKStep;
  const int64_t ri = a.L.row_of(i, rowok ? lr : 0), rj = a.L.row_of(j, rowok ? lr : 0);
  const float* pa = a.w32 + (ri / kGroupRows) * nkb * step + (ri % kGroupRows) * kKStep + 4 * lk;
  const float* pb = a.w32 + (rj / kGroupRows) * nkb * step + (rj % kGroupRows) * kKStep + 4 * lk;
  floatx4 acc = {0.f, 0.f, 0.f, 0.f};
  const floatx4 zero = {0.f, 0.f, 0.f, 0.f};
  for (int64_t s0 = 0; s0 < a.L.Sp; s0 += 16) {
    const int64_t o = (s0 / kKStep) * step + (s0 % kKStep);
    const floatx4 va = rowok ? *reinterpret_cast<const floatx4*>(pa + o) : zero;
    const floatx4 vb = rowok ? *reinterpret_cast<const floatx4*>(pb + o) : zero;
#pragma unroll
    for (int t = 0; t < 4; ++t) acc = __builtin_amdgcn_mfma_f32_16x16x4f32(va[t], vb[t], acc, 0, 0, 0);
  }
  const float m = a.m[n];
  double Ls = 0.0;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int b = 4 * lk + q, c = lr;
    if (b < bins && c < bins) {
      const float p = acc[q] * a.norm;
      float gm = 0.0f;
      if (p > 0.0f) {
        const float lp = log2f(p);
        Ls += (double)(p * lp);
        gm = m * (-(lp + 1.4426950408889634f));
      }
      a.Gm[n * bins * bins + b * bins + c] = gm;
    }
  }
  for (int o = 32; o > 0; o >>= 1) Ls += __shfl_xor(Ls, o);
  if (lane == 0) a.Lpart[n] = (double)m * ((a.H64[i] + a.H64[j]) + Ls);
}

struct GeneArgs {
  const float* w32;
  Layout L;
  Spline sp;
  const int* off;     // CSR offsets, G+1
  const int* nbr;     // CSR entries: (j, pair index, orientation) x nnz
  const double* cw;   // per gene: sum of m over its entries, in CSR order
  float* accg;        // MFMA path: neighbour sums, G x bins x Sp
  const float* Gm;
  const float* m;
  const double* pmarg;
  const float* xs;
  const float* scl;
  const int* amin;
  const int* amax;
  float* dX;
  int64_t lddx;
};

constexpr int kMaxSPerThread = 8;   // S <= 256 * 8 = 2048 samples in the adjoint

template <int NB, int NO>
__global__ __launch_bounds__(256) void bspline_gene_adjoint_kernel(GeneArgs a) {
  const int64_t i = blockIdx.x;
  const int tid = threadIdx.x;
  const int bins = a.L.bins;
  const int64_t S = a.L.S;
  __shared__ double shd[4];
  __shared__ float gsh[kMaxBins * kMaxBins];
  __shared__ float phi_i[kMaxBins];

  const int e0 = a.off[i], e1 = a.off[i + 1];
  // c_i * phi(p_i[b]) with c_i = sum of m over the gene's entries
  if (tid < bins) {
    const double p = a.pmarg[i * bins + tid];
    phi_i[tid] = p > 0.0 ? (float)(a.cw[i] * (-(log2(p) + 1.4426950408889634))) : 0.0f;
  }
  float accw[kMaxSPerThread][NB];
  const int nper = (int)ceil_div(S, 256);
#pragma unroll
  for (int q = 0; q < kMaxSPerThread; ++q)
#pragma unroll
    for (int b = 0; b < NB; ++b) accw[q][b] = 0.0f;

  for (int e = e0; e < e1; ++e) {
    const int j = a.nbr[3 * e + 0], n = a.nbr[3 * e + 1], ori = a.nbr[3 * e + 2];
    __syncthreads();
    for (int t = tid; t < bins * bins; t += 256) {
      const int b = t / bins, c = t % bins;
      // orientation 0: gene i is the row gene of the pair's table; 1: the column gene
      gsh[t] = ori == 0 ? a.Gm[(int64_t)n * bins * bins + b * bins + c] : a.Gm[(int64_t)n * bins * bins + c * bins + b];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < kMaxSPerThread; ++q) {
      const int64_t s = tid + 256 * q;
      if (q >= nper || s >= S) continue;
      float wj[NB];
#pragma unroll
      for (int c = 0; c < NB; ++c) wj[c] = c < bins ? a.w32[a.L.at(a.L.row_of(j, c), s)] : 0.0f;
#pragma unroll
      for (int b = 0; b < NB; ++b) {
        if (b < bins) {
          float v = 0.0f;
#pragma unroll
          for (int c = 0; c < NB; ++c)
            if (c < bins) v += gsh[b * bins + c] * wj[c];
          accw[q][b] += v;
        }
      }
    }
  }
  __syncthreads();
  // g_s = dL/dxhat_s = sum_b dL/dW[b][s] * N'_b(xhat_s),  dL/dW = (phi_i - acc) / S
  const float invS = 1.0f / (float)S;
  float gs[kMaxSPerThread];
  double s_lo = 0.0, s_hi = 0.0;
#pragma unroll
  for (int q = 0; q < kMaxSPerThread; ++q) {
    const int64_t s = tid + 256 * q;
    gs[q] = 0.0f;
    if (q >= nper || s >= S) continue;
    const float t = a.xs[i * S + s];
    float w[NB], lower[NB + 1], d[NB];
    basis_values<NB, NO>(a.sp, t, w, lower);
    basis_derivs<NB>(a.sp, lower, d);
    float g = 0.0f;
#pragma unroll
    for (int b = 0; b < NB; ++b)
      if (b < bins) g += (phi_i[b] - accw[q][b]) * invS * d[b];
    gs[q] = g;
    s_lo += (double)g * ((double)t - 1.0);
    s_hi += (double)g * (double)t;
  }
  const double sum_lo = block_sum<double>(s_lo, shd);
  const double sum_hi = block_sum<double>(s_hi, shd);
  const float scale = a.scl[i];
  const int imn = a.amin[i], imx = a.amax[i];
  const bool constant = !(scale != 1.0f || imn != imx);
#pragma unroll
  for (int q = 0; q < kMaxSPerThread; ++q) {
    const int64_t s = tid + 256 * q;
    if (q >= nper || s >= S) continue;
    double v = (double)gs[q] * (double)scale;
    if (!constant) {
      if (s == imn) v += sum_lo * (double)scale;
      if (s == imx) v -= sum_hi * (double)scale;
    }
    a.dX[i * a.lddx + s] = (float)v;
  }
}

// The same for bins <= 16 in two kernels.  bspline_gene_accumulate_kernel forms the neighbour
// sum acc_i[b][s] = sum_e G_e[b][c] W_j[c][s] on MFMA: a 16 x 16-sample output block of
// v_mfma_f32_16x16x4_f32 with K over c (four products; zero beyond bins).  Lane (r, k) supplies
// G_e[r][4t + k] and W_j[4t + k][s0 + r] in product t, and holds entry (4k + q, s0 + r) in output
// register q.  One workgroup per (256-sample pass, gene), 64 samples per wave as four output
// blocks whose sample r of block v is the wave's sample 4r + v (one float4 per W_j row and lane),
// through every neighbour; the sums go to a.accg (G x bins x Sp).  bspline_gene_epilogue_kernel then runs
// the per-sample derivative contraction.  Apart, the accumulation fits 60 VGPRs, eight waves per
// SIMD, and keeps more loads of the gathered W_j rows in flight (at 5000 x 500: 1.7 ms for both
// kernels; one kernel with its epilogue took 2.7 ms with four blocks per wave, 4.3 ms with eight).
constexpr int kAdjPassBlocks = 4;                        // 16-sample blocks per wave per pass
constexpr int kAdjPass = 4 * kAdjPassBlocks * 16;        // samples per pass (256)

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8, 8))) void bspline_gene_accumulate_kernel(GeneArgs a) {
  const int64_t i = blockIdx.y;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int lr = lane & 15, lk = lane >> 4;
  const int bins = a.L.bins;
  const int64_t Sp = a.L.Sp;
  const int64_t step = (int64_t)kGroupRows * kKStep;
  const int64_t nkb = Sp / kKStep;
  const int e0 = a.off[i], e1 = a.off[i + 1];
  // this wave's 64 samples: lane r holds samples s4 .. s4 + 3, sample s4 + v in output block v
  const int64_t s4 = (int64_t)blockIdx.x * kAdjPass + wave * (kAdjPassBlocks * 16) + 4 * lr;
  const bool sok = s4 < Sp;   // Sp is a multiple of 32, so s4 + 3 < Sp too
  const int64_t soff = sok ? (s4 / kKStep) * step + (s4 % kKStep) : 0;
  floatx4 acc[kAdjPassBlocks];
#pragma unroll
  for (int v = 0; v < kAdjPassBlocks; ++v) acc[v] = floatx4{0.f, 0.f, 0.f, 0.f};
  const floatx4 zero = {0.f, 0.f, 0.f, 0.f};
  // the next entry's (j, n, orientation) loads while this one's operands are used
  int jn = 0, nn = 0, on = 0;
  if (e0 < e1) { jn = a.nbr[3 * e0 + 0]; nn = a.nbr[3 * e0 + 1]; on = a.nbr[3 * e0 + 2]; }
  for (int e = e0; e < e1; ++e) {
    const int j = jn, n = nn, ori = on;
    if (e + 1 < e1) { jn = a.nbr[3 * e + 3]; nn = a.nbr[3 * e + 4]; on = a.nbr[3 * e + 5]; }
    const float* g = a.Gm + (int64_t)n * bins * bins;
    const int64_t rj = a.L.row_of(j, 0);
    const float* wj = a.w32 + (rj / kGroupRows) * nkb * step + (rj % kGroupRows) * kKStep + soff;
    float ga[4];
    floatx4 wv[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int c = 4 * t + lk;
      // orientation 0: gene i is the row gene of the pair's table; 1: the column gene
      ga[t] = (lr < bins && c < bins) ? (ori == 0 ? g[lr * bins + c] : g[c * bins + lr]) : 0.0f;
      wv[t] = (sok && c < bins) ? *reinterpret_cast<const floatx4*>(wj + c * kKStep) : zero;
    }
#pragma unroll
    for (int v = 0; v < kAdjPassBlocks; ++v)
#pragma unroll
      for (int t = 0; t < 4; ++t) acc[v] = __builtin_amdgcn_mfma_f32_16x16x4f32(ga[t], wv[t][v], acc[v], 0, 0, 0);
  }
  if (!sok) return;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int b = 4 * lk + q;
    if (b < bins)
      *reinterpret_cast<floatx4*>(a.accg + (i * bins + b) * Sp + s4) = floatx4{acc[0][q], acc[1][q], acc[2][q], acc[3][q]};
  }
}

__global__ __launch_bounds__(256) void bspline_gene_epilogue_kernel(GeneArgs a) {
  const int64_t i = blockIdx.x;
  const int tid = threadIdx.x;
  const int bins = a.L.bins;
  const int64_t S = a.L.S, Sp = a.L.Sp;
  __shared__ double shd[4];
  __shared__ float phi_i[16];
  __shared__ float gsh[kMaxSPerThread * 256];            // g_s for the whole row (S <= 2048)
  if (tid < bins) {
    const double p = a.pmarg[i * bins + tid];
    phi_i[tid] = p > 0.0 ? (float)(a.cw[i] * (-(log2(p) + 1.4426950408889634))) : 0.0f;
  }
  __syncthreads();
  // g_s = dL/dxhat_s = sum_b dL/dW[b][s] * N'_b(xhat_s),  dL/dW = (phi_i - acc) / S
  const float invS = 1.0f / (float)S;
  double s_lo = 0.0, s_hi = 0.0;
  for (int64_t s = tid; s < S; s += 256) {
    const float t = a.xs[i * S + s];
    float w[16], lower[17], d[16];
    basis_values<16, 4>(a.sp, t, w, lower);
    basis_derivs<16>(a.sp, lower, d);
    float gv = 0.0f;
#pragma unroll
    for (int b = 0; b < 16; ++b)
      if (b < bins) gv += (phi_i[b] - a.accg[(i * bins + b) * Sp + s]) * invS * d[b];
    gsh[s] = gv;
    s_lo += (double)gv * ((double)t - 1.0);
    s_hi += (double)gv * (double)t;
  }
  const double sum_lo = block_sum<double>(s_lo, shd);
  const double sum_hi = block_sum<double>(s_hi, shd);
  const float scale = a.scl[i];
  const int imn = a.amin[i], imx = a.amax[i];
  const bool constant = !(scale != 1.0f || imn != imx);
  for (int64_t s = tid; s < S; s += 256) {
    double v = (double)gsh[s] * (double)scale;
    if (!constant) {
      if (s == imn) v += sum_lo * (double)scale;
      if (s == imx) v -= sum_hi * (double)scale;
    }
    a.dX[i * a.lddx + s] = (float)v;
  }
}

// ---------------------------------------------------------------- host side
int check_spline_args(mcpg_ctx* h, int bins, int order) {
  if (order < 1 || order > kMaxOrder) return mcpg::set_error(h, MCPG_ERR_INVALID_ARG, "spline order must be in [1, 8]");
  if (bins < order) return mcpg::set_error(h, MCPG_ERR_INVALID_ARG, "bins must be >= spline order");
  if (bins > kMaxBins) return mcpg::set_error(h, MCPG_ERR_UNSUPPORTED, "bins must be <= 80");
  return MCPG_OK;
}

int prep_launch(mcpg_ctx* h, const float* dX, int64_t ldx, const Layout& L, const Spline& sp, int do_scale,
                const PrepOut& o) {
  mcpg::LaunchScope ls(h, "bspline_prep_kernel");
  if (o.whi_c && (L.bins != 10 || L.gw != 8 || (uint64_t)L.rows * (uint64_t)L.Sp * 2 >= (1ull << 32)))
    return mcpg::set_error(h, MCPG_ERR_INVALID_ARG, "internal: two-order planes need 10 bins and < 4 GiB");
  if (o.q8h && (L.ks != kKStep8 || (o.q8h_c && (L.bins != 10 || L.gw != 8))))
    return mcpg::set_error(h, MCPG_ERR_INVALID_ARG, "internal: int8 digit planes need 64-sample blocks (two orders: 10 bins)");
  // two-order planes: every gene slot of the planes, padding slots written as zeros
  const unsigned grid = (unsigned)((o.whi_c || o.q8h_c) ? L.ntiles * L.gt : L.G);
#define PREP_GO(NB, NO, TWO, I8)                                                                                 \
  hipLaunchKernelGGL((bspline_prep_kernel<NB, NO, TWO, I8>), dim3(grid), dim3(kThreads), 0, h->stream, dX, \
                     ldx, L, sp, do_scale, o)
  const bool small = sp.bins <= 16 && sp.order <= 4;
  if (o.q8h && o.q8h_c) {
    if (small) PREP_GO(16, 4, true, true);
    else PREP_GO(kMaxBins, kMaxOrder, true, true);
  } else if (o.q8h) {
    if (small) PREP_GO(16, 4, false, true);
    else PREP_GO(kMaxBins, kMaxOrder, false, true);
  } else if (o.whi_c) {
    if (small) PREP_GO(16, 4, true, false);
    else PREP_GO(kMaxBins, kMaxOrder, true, false);
  } else {
    if (small) PREP_GO(16, 4, false, false);
    else PREP_GO(kMaxBins, kMaxOrder, false, false);
  }
#undef PREP_GO
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return mcpg::hip_error(h, e, "bspline_prep_kernel launch");
  return MCPG_OK;
}

// BF16 planes in the two row orders of Layout::row_of_r / row_of_c (whi_c non-null) select the
// kernels with the register epilogue.
// A block of the G x G MI matrix (rows r0 .. r0 + nr - 1, columns c0 .. c0 + nc - 1) and the device
// list of the tiles that cover it; nr < 0 stands for the whole matrix.
struct MIBlock {
  int64_t r0 = 0, nr = -1, c0 = 0, nc = -1;
  const uint32_t* order = nullptr;
  int64_t ntiles = 0;
};

// The tiles (ti <= tj, packed (ti << 16) | tj) whose gene pairs land in the block at (i, j) or at
// (j, i), each once: every pair is computed in the orientation the whole-matrix launch uses, so a
// block holds exactly the whole matrix's values.  A tile position (ti, tj) with ti > tj is dropped
// when its mirror (tj, ti) is itself a position of the block.
void block_tiles(const Layout& L, int64_t r0, int64_t nr, int64_t c0, int64_t nc, std::vector<uint32_t>& out) {
  out.clear();
  if (nr <= 0 || nc <= 0) return;
  const int64_t gt = L.gt;
  const int64_t ra = r0 / gt, rb = (r0 + nr - 1) / gt, ca = c0 / gt, cb = (c0 + nc - 1) / gt;
  for (int64_t ti = ra; ti <= rb; ++ti)
    for (int64_t tj = ca; tj <= cb; ++tj) {
      if (ti > tj && tj >= ra && tj <= rb && ti >= ca && ti <= cb) continue;
      const int64_t lo = ti < tj ? ti : tj, hi = ti < tj ? tj : ti;
      out.push_back((uint32_t)((lo << 16) | hi));
    }
}

struct WeightBufs {
  float* w32 = nullptr;
  uint16_t* whi = nullptr;
  uint16_t* wlo = nullptr;
  uint16_t* whi_c = nullptr;   // row order C copies (two_orders)
  uint16_t* wlo_c = nullptr;
  int8_t* q8[4] = {nullptr, nullptr, nullptr, nullptr};   // I8X3 digit planes h, l (order R), h, l (order C)
  double* H64 = nullptr;
};

int mi_launch(mcpg_ctx* h, int precision, const WeightBufs& wb, const Layout& L, float* MI, int64_t ldmi,
              const MIBlock& blk = MIBlock()) {
  MIArgs a;
  a.w32 = wb.w32; a.whi = wb.whi; a.wlo = wb.wlo; a.whi_c = wb.whi_c; a.wlo_c = wb.wlo_c;
  for (int q = 0; q < 4; ++q) a.q8[q] = wb.q8[q];
  a.H64 = wb.H64; a.MI = MI; a.ldmi = ldmi; a.L = L;
  const bool whole = blk.nr < 0;
  a.r0 = whole ? 0 : blk.r0; a.nr = whole ? L.G : blk.nr;
  a.c0 = whole ? 0 : blk.c0; a.nc = whole ? L.G : blk.nc;
  const bool i8 = precision == MCPG_I8X3;
  const bool re = wb.whi_c != nullptr;
  const bool re8 = i8 && wb.q8[2] != nullptr;   // int8 planes in two row orders
  if (i8 && !re8) {   // one row order: operand B reads the operand-A planes
    a.q8[2] = wb.q8[0];
    a.q8[3] = wb.q8[1];
  }
  if (re && (precision == MCPG_FP32 || L.bins != 10))
    return mcpg::set_error(h, MCPG_ERR_INVALID_ARG, "internal: two-order planes need 10 bins and BF16");
  if (i8 && (!wb.q8[0] || L.ks != kKStep8))
    return mcpg::set_error(h, MCPG_ERR_INVALID_ARG, "internal: I8X3 needs the int8 digit planes");
  a.ntri = L.ntiles * (L.ntiles + 1) / 2;
  // p = acc * norm: acc sums w_i w_j (FP32 / BF16 kernels) or 127^2 w_i w_j (I8X3)
  a.norm = i8 ? (float)(1.0 / (16129.0 * (double)L.S)) : 1.0f / (float)L.S;
  if (a.ntri > (int64_t)INT32_MAX) return mcpg::set_error(h, MCPG_ERR_UNSUPPORTED, "MI grid too large");
  // the BF16 and int8 kernels address a weight plane through a buffer resource with a 32-bit byte range
  if (precision != MCPG_FP32 && (uint64_t)L.rows * (uint64_t)L.Sp * (i8 ? 1 : 2) >= (1ull << 32))
    return mcpg::set_error(h, MCPG_ERR_UNSUPPORTED, "weight panels exceed 4 GiB");
  int st;
  hipError_t e;
  if (L.tr != kTileRows) return mcpg::set_error(h, MCPG_ERR_INVALID_ARG, "internal: weight layout does not match the MI kernel");
  if (whole) {
    a.order = mcpg::tri_order(h, L.ntiles, mcpg::env_group("MCPG_MI_GROUP", 32), mcpg::env_group("MCPG_MI_XCD", 0), &st);
    if (!a.order) return st;
  } else {
    a.order = blk.order;
    a.ntri = blk.ntiles;
  }
  if (a.ntri == 0) return MCPG_OK;
  const unsigned grid = (unsigned)a.ntri;
  // BF16 planes: a ring of 2 (BF16X3, 80 KB of LDS) or 4 (BF16, 80 KB) stage buffers with early
  // release, two workgroups per CU; MCPG_MI_ER=0 selects the 2-stage ring without it
  const bool er = mcpg::env_group("MCPG_MI_ER", 1) != 0;
  static bool lds_attr_set = false;
  if (!lds_attr_set) {
#define MI_ATTR(P, N, E, R)                                                                               \
  MCPG_HIP(h, hipFuncSetAttribute(reinterpret_cast<const void*>(&bspline_mi_kernel<P, N, E, R>),         \
                                  hipFuncAttributeMaxDynamicSharedMemorySize, mi_smem_bytes<P, N>()))
    MI_ATTR(kFP32, 2, false, false);
    MI_ATTR(kBF16, 2, false, false);
    MI_ATTR(kBF16, 4, true, false);
    MI_ATTR(kBF16, 4, true, true);
    MI_ATTR(kBF16X3, 2, false, false);
    MI_ATTR(kBF16X3, 2, true, false);
    MI_ATTR(kBF16X3, 2, true, true);
#undef MI_ATTR
#define MI8_ATTR(N, E, R)                                                                                     \
  MCPG_HIP(h, hipFuncSetAttribute(reinterpret_cast<const void*>(&bspline_mi_i8_kernel<N, E, R>),             \
                                  hipFuncAttributeMaxDynamicSharedMemorySize, mi8_smem_bytes(N)))
    MI8_ATTR(2, false, true);
    MI8_ATTR(2, true, true);
    MI8_ATTR(3, false, true);
    MI8_ATTR(3, true, true);
    MI8_ATTR(2, false, false);
#undef MI8_ATTR
    lds_attr_set = true;
  }
  if (i8) {
    // MCPG_MI8: 0 = two stages, loads spread over the block rows (default); 1 = two stages, loads
    // issued together; 2, 3 = the same with three stages (120 KB, one workgroup per CU)
    const int v = mcpg::env_group("MCPG_MI8", 0);
    mcpg::LaunchScope ls(h, "bspline_mi_i8_kernel");
#define MI8_GO(N, E, R)                                                                                           \
  hipLaunchKernelGGL((bspline_mi_i8_kernel<N, E, R>), dim3(grid), dim3(kThreads), mi8_smem_bytes(N), h->stream, a)
    if (!re8) MI8_GO(2, false, false);
    else if (v == 1) MI8_GO(2, true, true);
    else if (v == 2) MI8_GO(3, false, true);
    else if (v == 3) MI8_GO(3, true, true);
    else MI8_GO(2, false, true);
#undef MI8_GO
  } else {
    const char* name = precision == MCPG_FP32 ? "bspline_mi_kernel<fp32>"
                       : precision == MCPG_BF16 ? "bspline_mi_kernel<bf16>" : "bspline_mi_kernel<bf16x3>";
    mcpg::LaunchScope ls(h, name);
#define MI_GO(P, N, E, R)                                                                                 \
  hipLaunchKernelGGL((bspline_mi_kernel<P, N, E, R>), dim3(grid), dim3(kThreads), (mi_smem_bytes<P, N>()), \
                     h->stream, a)
    if (re && !er) return mcpg::set_error(h, MCPG_ERR_INVALID_ARG, "internal: two-order planes need early release");
    if (precision == MCPG_FP32) MI_GO(kFP32, 2, false, false);
    else if (precision == MCPG_BF16 && re) MI_GO(kBF16, 4, true, true);
    else if (precision == MCPG_BF16 && er) MI_GO(kBF16, 4, true, false);
    else if (precision == MCPG_BF16) MI_GO(kBF16, 2, false, false);
    else if (re) MI_GO(kBF16X3, 2, true, true);
    else if (er) MI_GO(kBF16X3, 2, true, false);
    else MI_GO(kBF16X3, 2, false, false);
#undef MI_GO
  }
  e = hipGetLastError();
  if (e != hipSuccess) return mcpg::hip_error(h, e, "bspline_mi_kernel launch");
  return MCPG_OK;
}

// BF16 MI kernels with early release and 10 bins read their planes in two row orders and sum
// gene pairs in registers (MCPG_MI_ER=0 or MCPG_MI_RE=0 select the one-order planes).
bool two_orders(int precision, const Layout& L) {
  return (precision == MCPG_BF16 || precision == MCPG_BF16X3) && L.bins == 10 && L.gw == 8 &&
         (uint64_t)L.rows * (uint64_t)L.Sp * 2 < (1ull << 32) && mcpg::env_group("MCPG_MI_ER", 1) != 0 &&
         mcpg::env_group("MCPG_MI_RE", 1) != 0;
}

// int8 digit planes: two row orders and the register epilogue for 10 bins, else one row order
bool two_orders_i8(const Layout& L) { return L.bins == 10 && L.gw == 8; }

void size_weight_bufs(mcpg::Sizer& sz, const Layout& L, int precision) {
  const size_t n = (size_t)L.rows * L.Sp;
  const int copies = two_orders(precision, L) ? 2 : 1;
  if (precision == MCPG_FP32) sz.add<float>(n);
  else if (precision == MCPG_I8X3)
    for (int q = 0; q < (two_orders_i8(L) ? 4 : 2); ++q) sz.add<int8_t>(n);
  else
    for (int c = 0; c < copies; ++c) { sz.add<uint16_t>(n); if (precision == MCPG_BF16X3) sz.add<uint16_t>(n); }
  sz.add<double>(L.ntiles * L.gt);
}

WeightBufs take_weight_bufs(mcpg::Carver& cv, const Layout& L, int precision) {
  WeightBufs w;
  const size_t n = (size_t)L.rows * L.Sp;
  if (precision == MCPG_FP32) w.w32 = cv.take<float>(n);
  else if (precision == MCPG_I8X3) {
    for (int q = 0; q < (two_orders_i8(L) ? 4 : 2); ++q) w.q8[q] = cv.take<int8_t>(n);
  } else {
    w.whi = cv.take<uint16_t>(n);
    if (precision == MCPG_BF16X3) w.wlo = cv.take<uint16_t>(n);
    if (two_orders(precision, L)) {
      w.whi_c = cv.take<uint16_t>(n);
      if (precision == MCPG_BF16X3) w.wlo_c = cv.take<uint16_t>(n);
    }
  }
  w.H64 = cv.take<double>(L.ntiles * L.gt);
  return w;
}

int clear_weight_bufs(mcpg_ctx* h, const WeightBufs& w, const Layout& L) {
  const size_t n = (size_t)L.rows * L.Sp;
  if (w.w32) MCPG_HIP(h, hipMemsetAsync(w.w32, 0, n * sizeof(float), h->stream));
  for (uint16_t* p : {w.whi, w.wlo, w.whi_c, w.wlo_c})
    if (p) MCPG_HIP(h, hipMemsetAsync(p, 0, n * sizeof(uint16_t), h->stream));
  for (int8_t* p : w.q8)
    if (p) MCPG_HIP(h, hipMemsetAsync(p, 0, n, h->stream));
  return MCPG_OK;
}

int check_flags_nonfinite(mcpg_ctx* h, const char* what) {
  int flags[1];
  int st = mcpg::read_flags(h, flags, 1);
  if (st) return st;
  if (flags[0] & 1) return mcpg::set_error(h, MCPG_ERR_NONFINITE, std::string(what) + ": non-finite input value");
  return MCPG_OK;
}

bool valid_precision(int p) { return p == MCPG_FP32 || p == MCPG_BF16 || p == MCPG_BF16X3 || p == MCPG_I8X3; }

// gene groups per workgroup tile side of the MI kernel that runs at this precision
int mi_groups_per_tile(int) { return kTileRows / kGroupRows; }

// weight layout of the MI kernel that runs at this precision (64-sample K blocks for I8X3)
Layout mi_layout(int64_t G, int64_t S, int bins, int precision) {
  return make_layout(G, S, bins, mi_groups_per_tile(precision), precision == MCPG_I8X3 ? kKStep8 : kKStep);
}

// I8X3's int32 sums stay exact for padded sample counts with 2 * 127^2 * Sp < 2^31 (the largest
// 64-multiple is 66560).
int check_i8x3(mcpg_ctx* h, int precision, const Layout& L) {
  if (precision != MCPG_I8X3) return MCPG_OK;
  if (L.Sp > 66560) return mcpg::set_error(h, MCPG_ERR_UNSUPPORTED, "i8x3 precision supports at most 66560 samples");
  return MCPG_OK;
}

// compute_mi for one block of the output (the whole matrix: 0, G, 0, G): scaling, weights and
// marginals of every gene, then the MI launch over the tiles that cover the block.
int bspline_mi_run(mcpg_ctx* h, int ptrs, const float* X, int64_t G, int64_t S, int64_t ldx, int bins, int order,
                   int precision, int64_t r0, int64_t nr, int64_t c0, int64_t nc, float* MI, int64_t ldmi,
                   const char* where) {
  int st;
  if ((st = check_spline_args(h, bins, order))) return st;
  if (!valid_precision(precision)) return mcpg::set_error(h, MCPG_ERR_INVALID_ARG, "bad precision");
  if (G < 0 || S < 0) return mcpg::set_error(h, MCPG_ERR_SHAPE, "negative dimension");
  if (G == 0 || nr == 0 || nc == 0) return MCPG_OK;
  if (S < 1) return mcpg::set_error(h, MCPG_ERR_SHAPE, "need at least one sample");
  if (!X || !MI) return mcpg::set_error(h, MCPG_ERR_INVALID_ARG, "null matrix pointer");
  if (ldx < S || ldmi < nc) return mcpg::set_error(h, MCPG_ERR_INVALID_ARG, "leading dimension too small");
  if (ptrs != MCPG_HOST && ptrs != MCPG_DEVICE) return mcpg::set_error(h, MCPG_ERR_INVALID_ARG, "bad ptrs kind");
  MCPG_HIP(h, hipSetDevice(h->device));
  const Layout L = mi_layout(G, S, bins, precision);
  if ((st = check_i8x3(h, precision, L))) return st;
  const Spline sp = make_spline(bins, order);
  const bool whole = r0 == 0 && c0 == 0 && nr == G && nc == G;
  if (!whole) block_tiles(L, r0, nr, c0, nc, h->block_order);
  mcpg::Sizer sz;
  if (ptrs == MCPG_HOST) { sz.add<float>(G * S); sz.add<float>(nr * nc); }
  if (!whole) sz.add<uint32_t>(h->block_order.size());
  size_weight_bufs(sz, L, precision);
  void* base = mcpg::workspace(h, sz.off, &st);
  if (!base) return st;
  mcpg::Carver cv(base);
  const float* dX = X;
  int64_t lx = ldx;
  float* dMI = MI;
  int64_t lm = ldmi;
  if (ptrs == MCPG_HOST) {
    float* t = cv.take<float>(G * S);
    if ((st = mcpg::copy_2d(h, t, S * 4, X, ldx * 4, S * 4, G, hipMemcpyHostToDevice))) return st;
    dX = t; lx = S;
    dMI = cv.take<float>(nr * nc); lm = nc;
  }
  MIBlock blk;
  if (!whole) {
    uint32_t* d_order = cv.take<uint32_t>(h->block_order.size());
    MCPG_HIP(h, hipMemcpyAsync(d_order, h->block_order.data(), h->block_order.size() * 4, hipMemcpyHostToDevice,
                               h->stream));
    blk.r0 = r0; blk.nr = nr; blk.c0 = c0; blk.nc = nc;
    blk.order = d_order;
    blk.ntiles = (int64_t)h->block_order.size();
  }
  WeightBufs wb = take_weight_bufs(cv, L, precision);
  // the prep kernel writes every element of the two-order planes (BF16 and int8), padding included
  if (!wb.whi_c && !wb.q8[2] && (st = clear_weight_bufs(h, wb, L))) return st;
  if ((st = mcpg::reset_flags(h))) return st;
  PrepOut o{};
  o.w32 = wb.w32; o.whi = wb.whi; o.wlo = wb.wlo; o.whi_c = wb.whi_c; o.wlo_c = wb.wlo_c;
  o.q8h = wb.q8[0]; o.q8l = wb.q8[1]; o.q8h_c = wb.q8[2]; o.q8l_c = wb.q8[3];
  o.H64 = wb.H64; o.flags = h->d_flags;
  if ((st = prep_launch(h, dX, lx, L, sp, 1, o))) return st;
  if ((st = mi_launch(h, precision, wb, L, dMI, lm, blk))) return st;
  if (ptrs == MCPG_HOST && (st = mcpg::copy_2d(h, MI, ldmi * 4, dMI, nc * 4, nc * 4, nr, hipMemcpyDeviceToHost)))
    return st;
  if ((st = mcpg::finish(h, where))) return st;
  return check_flags_nonfinite(h, where);
}

// ---- FP64 host side
Layout mi64_layout(int64_t G, int64_t S, int bins) { return make_layout(G, S, bins, 2, kKStep64); }

int prep64_launch(mcpg_ctx* h, const double* dX, int64_t ldx, const Layout& L, const Spline64& sp, int do_scale,
                  const Prep64Out& o) {
  mcpg::LaunchScope ls(h, "bspline_prep64_kernel");
  if (sp.bins <= 16 && sp.order <= 4)
    hipLaunchKernelGGL((bspline_prep64_kernel<16, 4>), dim3((unsigned)L.G), dim3(kThreads), 0, h->stream, dX, ldx, L,
                       sp, do_scale, o);
  else
    hipLaunchKernelGGL((bspline_prep64_kernel<kMaxBins, kMaxOrder>), dim3((unsigned)L.G), dim3(kThreads), 0,
                       h->stream, dX, ldx, L, sp, do_scale, o);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return mcpg::hip_error(h, e, "bspline_prep64_kernel launch");
  return MCPG_OK;
}

int mi64_launch(mcpg_ctx* h, const double* w, const double* H64, const Layout& L, double* MI, int64_t ldmi,
                const MIBlock& blk = MIBlock()) {
  if (L.tr != kTileRows || L.ks != kKStep64)
    return mcpg::set_error(h, MCPG_ERR_INVALID_ARG, "internal: weight layout does not match the FP64 MI kernel");
  const bool whole = blk.nr < 0;
  int64_t ntri = L.ntiles * (L.ntiles + 1) / 2;
  if (ntri > (int64_t)INT32_MAX || L.ntiles > 65535) return mcpg::set_error(h, MCPG_ERR_UNSUPPORTED, "MI grid too large");
  MI64Args a;
  a.w = w; a.H64 = H64; a.MI = MI; a.ldmi = ldmi; a.L = L;
  a.norm = 1.0 / (double)L.S;
  a.r0 = whole ? 0 : blk.r0; a.nr = whole ? L.G : blk.nr;
  a.c0 = whole ? 0 : blk.c0; a.nc = whole ? L.G : blk.nc;
  int st;
  if (whole) {
    a.order = mcpg::tri_order(h, L.ntiles, mcpg::env_group("MCPG_MI_GROUP", 32), mcpg::env_group("MCPG_MI_XCD", 0), &st);
    if (!a.order) return st;
  } else {
    a.order = blk.order;
    ntri = blk.ntiles;
  }
  if (ntri == 0) return MCPG_OK;
  static bool attr_set = false;
  if (!attr_set) {
    MCPG_HIP(h, hipFuncSetAttribute(reinterpret_cast<const void*>(&bspline_mi64_kernel<2>),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, kEpi64Bytes));
    MCPG_HIP(h, hipFuncSetAttribute(reinterpret_cast<const void*>(&bspline_mi64_kernel<3>),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, kEpi64Bytes));
    MCPG_HIP(h, hipFuncSetAttribute(reinterpret_cast<const void*>(&bspline_mi64_kernel<4>),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, kEpi64Bytes));
    attr_set = true;
  }
  // MCPG_MI64_BUF: register sets in the load ring (2 to 4; 3 by default)
  const int nbuf = std::min(std::max(mcpg::env_group("MCPG_MI64_BUF", 3), 2), 4);
  mcpg::LaunchScope ls(h, "bspline_mi64_kernel");
  if (nbuf == 2)
    hipLaunchKernelGGL(bspline_mi64_kernel<2>, dim3((unsigned)ntri), dim3(kThreads), kEpi64Bytes, h->stream, a);
  else if (nbuf == 4)
    hipLaunchKernelGGL(bspline_mi64_kernel<4>, dim3((unsigned)ntri), dim3(kThreads), kEpi64Bytes, h->stream, a);
  else
    hipLaunchKernelGGL(bspline_mi64_kernel<3>, dim3((unsigned)ntri), dim3(kThreads), kEpi64Bytes, h->stream, a);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return mcpg::hip_error(h, e, "bspline_mi64_kernel launch");
  return MCPG_OK;
}

// compute_mi on the reference's double DataType for one block of the output (the whole matrix:
// 0, G, 0, G): scaling, weights and marginals of every gene in double, then the FP64 MI of the
// tiles that cover the block.
int bspline_mi64_run(mcpg_ctx* h, int ptrs, const double* X, int64_t G, int64_t S, int64_t ldx, int bins, int order,
                     int64_t r0, int64_t nr, int64_t c0, int64_t nc, double* MI, int64_t ldmi, const char* where) {
  int st;
  if ((st = check_spline_args(h, bins, order))) return st;
  if (G < 0 || S < 0) return mcpg::set_error(h, MCPG_ERR_SHAPE, "negative dimension");
  if (G == 0 || nr == 0 || nc == 0) return MCPG_OK;
  if (S < 1) return mcpg::set_error(h, MCPG_ERR_SHAPE, "need at least one sample");
  if (!X || !MI) return mcpg::set_error(h, MCPG_ERR_INVALID_ARG, "null matrix pointer");
  if (ldx < S || ldmi < nc) return mcpg::set_error(h, MCPG_ERR_INVALID_ARG, "leading dimension too small");
  if (ptrs != MCPG_HOST && ptrs != MCPG_DEVICE) return mcpg::set_error(h, MCPG_ERR_INVALID_ARG, "bad ptrs kind");
  MCPG_HIP(h, hipSetDevice(h->device));
  const Layout L = mi64_layout(G, S, bins);
  const Spline64 sp = make_spline64(bins, order);
  const bool whole = r0 == 0 && c0 == 0 && nr == G && nc == G;
  if (!whole) block_tiles(L, r0, nr, c0, nc, h->block_order);
  const size_t n = (size_t)L.rows * L.Sp;
  mcpg::Sizer sz;
  if (ptrs == MCPG_HOST) { sz.add<double>(G * S); sz.add<double>(nr * nc); }
  if (!whole) sz.add<uint32_t>(h->block_order.size());
  sz.add<double>(n);
  sz.add<double>(L.ntiles * L.gt);
  void* base = mcpg::workspace(h, sz.off, &st);
  if (!base) return st;
  mcpg::Carver cv(base);
  const double* dX = X;
  int64_t lx = ldx;
  double* dMI = MI;
  int64_t lm = ldmi;
  if (ptrs == MCPG_HOST) {
    double* t = cv.take<double>(G * S);
    if ((st = mcpg::copy_2d(h, t, S * 8, X, ldx * 8, S * 8, G, hipMemcpyHostToDevice))) return st;
    dX = t; lx = S;
    dMI = cv.take<double>(nr * nc); lm = nc;
  }
  MIBlock blk;
  if (!whole) {
    uint32_t* d_order = cv.take<uint32_t>(h->block_order.size());
    MCPG_HIP(h, hipMemcpyAsync(d_order, h->block_order.data(), h->block_order.size() * 4, hipMemcpyHostToDevice,
                               h->stream));
    blk.r0 = r0; blk.nr = nr; blk.c0 = c0; blk.nc = nc;
    blk.order = d_order;
    blk.ntiles = (int64_t)h->block_order.size();
  }
  double* w = cv.take<double>(n);
  double* H64 = cv.take<double>(L.ntiles * L.gt);
  MCPG_HIP(h, hipMemsetAsync(w, 0, n * sizeof(double), h->stream));   // padding rows and samples
  if ((st = mcpg::reset_flags(h))) return st;
  Prep64Out o{};
  o.w = w; o.H64 = H64; o.flags = h->d_flags;
  if ((st = prep64_launch(h, dX, lx, L, sp, 1, o))) return st;
  if ((st = mi64_launch(h, w, H64, L, dMI, lm, blk))) return st;
  if (ptrs == MCPG_HOST && (st = mcpg::copy_2d(h, MI, ldmi * 8, dMI, nc * 8, nc * 8, nr, hipMemcpyDeviceToHost)))
    return st;
  if ((st = mcpg::finish(h, where))) return st;
  return check_flags_nonfinite(h, where);
}

}  // namespace

extern "C" {

int mcpg_minmax_scale_f32(mcpg_handle h, int ptrs, const float* X, int64_t G, int64_t S, int64_t ldx, float* Y,
                          int64_t ldy) {
  MCPG_CHECK_HANDLE(h);
  mcpg::begin_call(h);
  if (G < 0 || S < 0) return mcpg::set_error(h, MCPG_ERR_SHAPE, "negative dimension");
  if (G == 0 || S == 0) return MCPG_OK;
  if (!X || !Y) return mcpg::set_error(h, MCPG_ERR_INVALID_ARG, "null matrix pointer");
  if (ldx < S || ldy < S) return mcpg::set_error(h, MCPG_ERR_INVALID_ARG, "leading dimension too small");
  if (ptrs != MCPG_HOST && ptrs != MCPG_DEVICE) return mcpg::set_error(h, MCPG_ERR_INVALID_ARG, "bad ptrs kind");
  MCPG_HIP(h, hipSetDevice(h->device));
  const Layout L = make_layout(G, S, 10);
  const Spline sp = make_spline(10, 3);
  mcpg::Sizer sz;
  if (ptrs == MCPG_HOST) sz.add<float>(G * S);
  sz.add<float>(G * S);
  int st;
  void* base = mcpg::workspace(h, sz.off, &st);
  if (!base) return st;
  mcpg::Carver cv(base);
  const float* dX = X;
  int64_t lx = ldx;
  if (ptrs == MCPG_HOST) {
    float* t = cv.take<float>(G * S);
    if ((st = mcpg::copy_2d(h, t, S * 4, X, ldx * 4, S * 4, G, hipMemcpyHostToDevice))) return st;
    dX = t; lx = S;
  }
  float* xs = cv.take<float>(G * S);
  if ((st = mcpg::reset_flags(h))) return st;
  PrepOut o{};
  o.xs = xs; o.flags = h->d_flags;
  if ((st = prep_launch(h, dX, lx, L, sp, 1, o))) return st;
  if (ptrs == MCPG_HOST) {
    if ((st = mcpg::copy_2d(h, Y, ldy * 4, xs, S * 4, S * 4, G, hipMemcpyDeviceToHost))) return st;
  } else {
    if ((st = mcpg::copy_2d(h, Y, ldy * 4, xs, S * 4, S * 4, G, hipMemcpyDeviceToDevice))) return st;
  }
  if ((st = mcpg::finish(h, "mcpg_minmax_scale_f32"))) return st;
  return check_flags_nonfinite(h, "mcpg_minmax_scale_f32");
}

int mcpg_bspline_weights_f32(mcpg_handle h, int ptrs, const float* Xs, int64_t G, int64_t S, int64_t ldx, int bins,
                             int order, float* W, int64_t ldw) {
  MCPG_CHECK_HANDLE(h);
  mcpg::begin_call(h);
  int st;
  if ((st = check_spline_args(h, bins, order))) return st;
  if (G < 0 || S < 0) return mcpg::set_error(h, MCPG_ERR_SHAPE, "negative dimension");
  if (G == 0) return MCPG_OK;
  if (S < 1) return mcpg::set_error(h, MCPG_ERR_SHAPE, "need at least one sample");
  if (!Xs || !W) return mcpg::set_error(h, MCPG_ERR_INVALID_ARG, "null matrix pointer");
  if (ldx < S || ldw < bins * S + 1) return mcpg::set_error(h, MCPG_ERR_INVALID_ARG, "leading dimension too small");
  if (ptrs != MCPG_HOST && ptrs != MCPG_DEVICE) return mcpg::set_error(h, MCPG_ERR_INVALID_ARG, "bad ptrs kind");
  MCPG_HIP(h, hipSetDevice(h->device));
  const Layout L = make_layout(G, S, bins);
  const Spline sp = make_spline(bins, order);
  const int64_t wrow = (int64_t)bins * S + 1;
  mcpg::Sizer sz;
  if (ptrs == MCPG_HOST) { sz.add<float>(G * S); sz.add<float>(G * wrow); }
  int64_t lw = ldw;
  void* base = sz.off ? mcpg::workspace(h, sz.off, &st) : h->ws;
  if (sz.off && !base) return st;
  mcpg::Carver cv(base);
  const float* dX = Xs;
  int64_t lx = ldx;
  float* dW = W;
  if (ptrs == MCPG_HOST) {
    float* t = cv.take<float>(G * S);
    if ((st = mcpg::copy_2d(h, t, S * 4, Xs, ldx * 4, S * 4, G, hipMemcpyHostToDevice))) return st;
    dX = t; lx = S;
    dW = cv.take<float>(G * wrow);
    lw = wrow;
  }
  if ((st = mcpg::reset_flags(h))) return st;
  PrepOut o{};
  o.wref = dW; o.ldw = lw; o.flags = h->d_flags;
  if ((st = prep_launch(h, dX, lx, L, sp, 0, o))) return st;
  if (ptrs == MCPG_HOST &&
      (st = mcpg::copy_2d(h, W, ldw * 4, dW, wrow * 4, wrow * 4, G, hipMemcpyDeviceToHost)))
    return st;
  if ((st = mcpg::finish(h, "mcpg_bspline_weights_f32"))) return st;
  return check_flags_nonfinite(h, "mcpg_bspline_weights_f32");
}

int mcpg_bspline_mi_weights_f32(mcpg_handle h, int ptrs, const float* W, int64_t G, int64_t S, int64_t ldw, int bins,
                                int precision, float* MI, int64_t ldmi) {
  MCPG_CHECK_HANDLE(h);
  mcpg::begin_call(h);
  if (bins < 1 || bins > kMaxBins) return mcpg::set_error(h, MCPG_ERR_INVALID_ARG, "bins must be in [1, 80]");
  if (!valid_precision(precision)) return mcpg::set_error(h, MCPG_ERR_INVALID_ARG, "bad precision");
  if (G < 0 || S < 0) return mcpg::set_error(h, MCPG_ERR_SHAPE, "negative dimension");
  if (G == 0) return MCPG_OK;
  if (S < 1) return mcpg::set_error(h, MCPG_ERR_SHAPE, "need at least one sample");
  if (!W || !MI) return mcpg::set_error(h, MCPG_ERR_INVALID_ARG, "null matrix pointer");
  if (ldw < bins * S + 1 || ldmi < G) return mcpg::set_error(h, MCPG_ERR_SHAPE, "weight rows must hold bins*S+1 values");
  if (ptrs != MCPG_HOST && ptrs != MCPG_DEVICE) return mcpg::set_error(h, MCPG_ERR_INVALID_ARG, "bad ptrs kind");
  MCPG_HIP(h, hipSetDevice(h->device));
  const Layout L = mi_layout(G, S, bins, precision);
  int st;
  if ((st = check_i8x3(h, precision, L))) return st;
  const int64_t wrow = (int64_t)bins * S + 1;
  mcpg::Sizer sz;
  if (ptrs == MCPG_HOST) { sz.add<float>(G * wrow); sz.add<float>(G * G); }
  size_weight_bufs(sz, L, precision);
  void* base = mcpg::workspace(h, sz.off, &st);
  if (!base) return st;
  mcpg::Carver cv(base);
  const float* dW = W;
  int64_t lw = ldw;
  float* dMI = MI;
  int64_t lm = ldmi;
  if (ptrs == MCPG_HOST) {
    float* t = cv.take<float>(G * wrow);
    if ((st = mcpg::copy_2d(h, t, wrow * 4, W, ldw * 4, wrow * 4, G, hipMemcpyHostToDevice))) return st;
    dW = t; lw = wrow;
    dMI = cv.take<float>(G * G); lm = G;
  }
  WeightBufs wb = take_weight_bufs(cv, L, precision);
  if ((st = clear_weight_bufs(h, wb, L))) return st;
  if ((st = mcpg::reset_flags(h))) return st;
  {
    mcpg::LaunchScope ls(h, "weights_import_kernel");
    hipLaunchKernelGGL(weights_import_kernel, dim3((unsigned)std::min<int64_t>(ceil_div(wrow, 256), 64), (unsigned)G),
                       dim3(256), 0, h->stream, dW, lw, L, wb.w32, wb.whi, wb.wlo, wb.whi_c, wb.wlo_c, wb.q8[0],
                       wb.q8[1], wb.q8[2], wb.q8[3], wb.H64, h->d_flags);
  }
  if ((st = mi_launch(h, precision, wb, L, dMI, lm))) return st;
  if (ptrs == MCPG_HOST && (st = mcpg::copy_2d(h, MI, ldmi * 4, dMI, G * 4, G * 4, G, hipMemcpyDeviceToHost))) return st;
  if ((st = mcpg::finish(h, "mcpg_bspline_mi_weights_f32"))) return st;
  int flags[1];
  if ((st = mcpg::read_flags(h, flags, 1))) return st;
  if (flags[0] & 1) return mcpg::set_error(h, MCPG_ERR_NONFINITE, "mcpg_bspline_mi_weights_f32: non-finite input value");
  if (flags[0] & 2)
    return mcpg::set_error(h, MCPG_ERR_INVALID_ARG, "mcpg_bspline_mi_weights_f32: i8x3 needs weights in [0, 1]");
  return MCPG_OK;
}

int mcpg_bspline_mi_f32(mcpg_handle h, int ptrs, const float* X, int64_t G, int64_t S, int64_t ldx, int bins, int order,
                        int precision, float* MI, int64_t ldmi) {
  MCPG_CHECK_HANDLE(h);
  mcpg::begin_call(h);
  return bspline_mi_run(h, ptrs, X, G, S, ldx, bins, order, precision, 0, G, 0, G, MI, ldmi, "mcpg_bspline_mi_f32");
}

int mcpg_bspline_mi_block_f32(mcpg_handle h, int ptrs, const float* X, int64_t G, int64_t S, int64_t ldx, int bins,
                              int order, int precision, int64_t row0, int64_t nrows, int64_t col0, int64_t ncols,
                              float* MI, int64_t ldmi) {
  MCPG_CHECK_HANDLE(h);
  mcpg::begin_call(h);
  if (G < 0 || row0 < 0 || nrows < 0 || col0 < 0 || ncols < 0 || row0 > G - nrows || col0 > G - ncols)
    return mcpg::set_error(h, MCPG_ERR_SHAPE, "block outside the G x G matrix");
  return bspline_mi_run(h, ptrs, X, G, S, ldx, bins, order, precision, row0, nrows, col0, ncols, MI, ldmi,
                        "mcpg_bspline_mi_block_f32");
}

int mcpg_minmax_scale_f64(mcpg_handle h, int ptrs, const double* X, int64_t G, int64_t S, int64_t ldx, double* Y,
                          int64_t ldy) {
  MCPG_CHECK_HANDLE(h);
  mcpg::begin_call(h);
  if (G < 0 || S < 0) return mcpg::set_error(h, MCPG_ERR_SHAPE, "negative dimension");
  if (G == 0 || S == 0) return MCPG_OK;
  if (!X || !Y) return mcpg::set_error(h, MCPG_ERR_INVALID_ARG, "null matrix pointer");
  if (ldx < S || ldy < S) return mcpg::set_error(h, MCPG_ERR_INVALID_ARG, "leading dimension too small");
  if (ptrs != MCPG_HOST && ptrs != MCPG_DEVICE) return mcpg::set_error(h, MCPG_ERR_INVALID_ARG, "bad ptrs kind");
  MCPG_HIP(h, hipSetDevice(h->device));
  const Layout L = mi64_layout(G, S, 10);
  const Spline64 sp = make_spline64(10, 3);
  mcpg::Sizer sz;
  if (ptrs == MCPG_HOST) { sz.add<double>(G * S); sz.add<double>(G * S); }
  int st;
  void* base = sz.off ? mcpg::workspace(h, sz.off, &st) : h->ws;
  if (sz.off && !base) return st;
  mcpg::Carver cv(base);
  const double* dX = X;
  int64_t lx = ldx;
  double* dY = Y;
  int64_t ly = ldy;
  if (ptrs == MCPG_HOST) {
    double* t = cv.take<double>(G * S);
    if ((st = mcpg::copy_2d(h, t, S * 8, X, ldx * 8, S * 8, G, hipMemcpyHostToDevice))) return st;
    dX = t; lx = S;
    dY = cv.take<double>(G * S); ly = S;
  }
  if ((st = mcpg::reset_flags(h))) return st;
  Prep64Out o{};
  o.xs = dY; o.ldxs = ly; o.flags = h->d_flags;
  if ((st = prep64_launch(h, dX, lx, L, sp, 1, o))) return st;
  if (ptrs == MCPG_HOST && (st = mcpg::copy_2d(h, Y, ldy * 8, dY, S * 8, S * 8, G, hipMemcpyDeviceToHost))) return st;
  if ((st = mcpg::finish(h, "mcpg_minmax_scale_f64"))) return st;
  return check_flags_nonfinite(h, "mcpg_minmax_scale_f64");
}

int mcpg_bspline_weights_f64(mcpg_handle h, int ptrs, const double* Xs, int64_t G, int64_t S, int64_t ldx, int bins,
                             int order, double* W, int64_t ldw) {
  MCPG_CHECK_HANDLE(h);
  mcpg::begin_call(h);
  int st;
  if ((st = check_spline_args(h, bins, order))) return st;
  if (G < 0 || S < 0) return mcpg::set_error(h, MCPG_ERR_SHAPE, "negative dimension");
  if (G == 0) return MCPG_OK;
  if (S < 1) return mcpg::set_error(h, MCPG_ERR_SHAPE, "need at least one sample");
  if (!Xs || !W) return mcpg::set_error(h, MCPG_ERR_INVALID_ARG, "null matrix pointer");
  if (ldx < S || ldw < bins * S + 1) return mcpg::set_error(h, MCPG_ERR_INVALID_ARG, "leading dimension too small");
  if (ptrs != MCPG_HOST && ptrs != MCPG_DEVICE) return mcpg::set_error(h, MCPG_ERR_INVALID_ARG, "bad ptrs kind");
  MCPG_HIP(h, hipSetDevice(h->device));
  const Layout L = mi64_layout(G, S, bins);
  const Spline64 sp = make_spline64(bins, order);
  const int64_t wrow = (int64_t)bins * S + 1;
  mcpg::Sizer sz;
  if (ptrs == MCPG_HOST) { sz.add<double>(G * S); sz.add<double>(G * wrow); }
  void* base = sz.off ? mcpg::workspace(h, sz.off, &st) : h->ws;
  if (sz.off && !base) return st;
  mcpg::Carver cv(base);
  const double* dX = Xs;
  int64_t lx = ldx;
  double* dW = W;
  int64_t lw = ldw;
  if (ptrs == MCPG_HOST) {
    double* t = cv.take<double>(G * S);
    if ((st = mcpg::copy_2d(h, t, S * 8, Xs, ldx * 8, S * 8, G, hipMemcpyHostToDevice))) return st;
    dX = t; lx = S;
    dW = cv.take<double>(G * wrow); lw = wrow;
  }
  if ((st = mcpg::reset_flags(h))) return st;
  Prep64Out o{};
  o.wref = dW; o.ldw = lw; o.flags = h->d_flags;
  if ((st = prep64_launch(h, dX, lx, L, sp, 0, o))) return st;
  if (ptrs == MCPG_HOST && (st = mcpg::copy_2d(h, W, ldw * 8, dW, wrow * 8, wrow * 8, G, hipMemcpyDeviceToHost)))
    return st;
  if ((st = mcpg::finish(h, "mcpg_bspline_weights_f64"))) return st;
  return check_flags_nonfinite(h, "mcpg_bspline_weights_f64");
}

int mcpg_bspline_mi_weights_f64(mcpg_handle h, int ptrs, const double* W, int64_t G, int64_t S, int64_t ldw, int bins,
                                double* MI, int64_t ldmi) {
  MCPG_CHECK_HANDLE(h);
  mcpg::begin_call(h);
  if (bins < 1 || bins > kMaxBins) return mcpg::set_error(h, MCPG_ERR_INVALID_ARG, "bins must be in [1, 80]");
  if (G < 0 || S < 0) return mcpg::set_error(h, MCPG_ERR_SHAPE, "negative dimension");
  if (G == 0) return MCPG_OK;
  if (S < 1) return mcpg::set_error(h, MCPG_ERR_SHAPE, "need at least one sample");
  if (!W || !MI) return mcpg::set_error(h, MCPG_ERR_INVALID_ARG, "null matrix pointer");
  if (ldw < bins * S + 1 || ldmi < G) return mcpg::set_error(h, MCPG_ERR_SHAPE, "weight rows must hold bins*S+1 values");
  if (ptrs != MCPG_HOST && ptrs != MCPG_DEVICE) return mcpg::set_error(h, MCPG_ERR_INVALID_ARG, "bad ptrs kind");
  MCPG_HIP(h, hipSetDevice(h->device));
  const Layout L = mi64_layout(G, S, bins);
  const int64_t wrow = (int64_t)bins * S + 1;
  const size_t n = (size_t)L.rows * L.Sp;
  int st;
  mcpg::Sizer sz;
  if (ptrs == MCPG_HOST) { sz.add<double>(G * wrow); sz.add<double>(G * G); }
  sz.add<double>(n);
  sz.add<double>(L.ntiles * L.gt);
  void* base = mcpg::workspace(h, sz.off, &st);
  if (!base) return st;
  mcpg::Carver cv(base);
  const double* dW = W;
  int64_t lw = ldw;
  double* dMI = MI;
  int64_t lm = ldmi;
  if (ptrs == MCPG_HOST) {
    double* t = cv.take<double>(G * wrow);
    if ((st = mcpg::copy_2d(h, t, wrow * 8, W, ldw * 8, wrow * 8, G, hipMemcpyHostToDevice))) return st;
    dW = t; lw = wrow;
    dMI = cv.take<double>(G * G); lm = G;
  }
  double* w = cv.take<double>(n);
  double* H64 = cv.take<double>(L.ntiles * L.gt);
  MCPG_HIP(h, hipMemsetAsync(w, 0, n * sizeof(double), h->stream));
  if ((st = mcpg::reset_flags(h))) return st;
  {
    mcpg::LaunchScope ls(h, "weights_import64_kernel");
    hipLaunchKernelGGL(weights_import64_kernel,
                       dim3((unsigned)std::min<int64_t>(ceil_div(wrow, 256), 64), (unsigned)std::min<int64_t>(G, 65535)),
                       dim3(256), 0, h->stream, dW, lw, L, w, H64, h->d_flags);
  }
  if ((st = mi64_launch(h, w, H64, L, dMI, lm))) return st;
  if (ptrs == MCPG_HOST && (st = mcpg::copy_2d(h, MI, ldmi * 8, dMI, G * 8, G * 8, G, hipMemcpyDeviceToHost))) return st;
  if ((st = mcpg::finish(h, "mcpg_bspline_mi_weights_f64"))) return st;
  return check_flags_nonfinite(h, "mcpg_bspline_mi_weights_f64");
}

int mcpg_bspline_mi_f64(mcpg_handle h, int ptrs, const double* X, int64_t G, int64_t S, int64_t ldx, int bins, int order,
                        double* MI, int64_t ldmi) {
  MCPG_CHECK_HANDLE(h);
  mcpg::begin_call(h);
  return bspline_mi64_run(h, ptrs, X, G, S, ldx, bins, order, 0, G, 0, G, MI, ldmi, "mcpg_bspline_mi_f64");
}

int mcpg_bspline_mi_block_f64(mcpg_handle h, int ptrs, const double* X, int64_t G, int64_t S, int64_t ldx, int bins,
                              int order, int64_t row0, int64_t nrows, int64_t col0, int64_t ncols, double* MI,
                              int64_t ldmi) {
  MCPG_CHECK_HANDLE(h);
  mcpg::begin_call(h);
  if (G < 0 || row0 < 0 || nrows < 0 || col0 < 0 || ncols < 0 || row0 > G - nrows || col0 > G - ncols)
    return mcpg::set_error(h, MCPG_ERR_SHAPE, "block outside the G x G matrix");
  return bspline_mi64_run(h, ptrs, X, G, S, ldx, bins, order, row0, nrows, col0, ncols, MI, ldmi,
                          "mcpg_bspline_mi_block_f64");
}

int mcpg_bspline_mi_backward_f32(mcpg_handle h, int ptrs, const float* X, int64_t G, int64_t S, int64_t ldx, int bins,
                                 int order, const int32_t* pi, const int32_t* pj, const float* m, int64_t npairs,
                                 float* dX, int64_t lddx, double* L_out) {
  MCPG_CHECK_HANDLE(h);
  mcpg::begin_call(h);
  int st;
  if ((st = check_spline_args(h, bins, order))) return st;
  if (G < 0 || S < 0 || npairs < 0) return mcpg::set_error(h, MCPG_ERR_SHAPE, "negative dimension");
  if (G == 0) return MCPG_OK;
  if (S < 1) return mcpg::set_error(h, MCPG_ERR_SHAPE, "need at least one sample");
  if (S > 256 * kMaxSPerThread) return mcpg::set_error(h, MCPG_ERR_UNSUPPORTED, "adjoint supports S <= 2048");
  if (bins > 16 || order > 4) return mcpg::set_error(h, MCPG_ERR_UNSUPPORTED, "adjoint supports bins <= 16, order <= 4");
  if (!X || !dX || (npairs > 0 && (!pi || !pj || !m))) return mcpg::set_error(h, MCPG_ERR_INVALID_ARG, "null pointer");
  if (ldx < S || lddx < S) return mcpg::set_error(h, MCPG_ERR_INVALID_ARG, "leading dimension too small");
  if (ptrs != MCPG_HOST && ptrs != MCPG_DEVICE) return mcpg::set_error(h, MCPG_ERR_INVALID_ARG, "bad ptrs kind");
  if (npairs > (int64_t)INT32_MAX / 2) return mcpg::set_error(h, MCPG_ERR_UNSUPPORTED, "too many pairs");
  // CSR of the symmetric pair list: gene -> (other gene, pair index, orientation); the counts
  // first, the entries below, written straight into the block that goes to the device
  std::vector<int> off(G + 1, 0);
  for (int64_t n = 0; n < npairs; ++n) {
    if (pi[n] < 0 || pi[n] >= G || pj[n] < 0 || pj[n] >= G)
      return mcpg::set_error(h, MCPG_ERR_INVALID_ARG, "pair index out of range");
    off[pi[n] + 1]++;
    off[pj[n] + 1]++;
  }
  for (int64_t g = 0; g < G; ++g) off[g + 1] += off[g];
  const size_t nnbr = 3 * (size_t)off[G] + 3;
  MCPG_HIP(h, hipSetDevice(h->device));
  const Layout L = make_layout(G, S, bins);
  const Spline sp = make_spline(bins, order);
  const int64_t np1 = npairs > 0 ? npairs : 1;
  mcpg::Sizer sz;
  if (ptrs == MCPG_HOST) { sz.add<float>(G * S); sz.add<float>(G * S); }
  size_weight_bufs(sz, L, MCPG_FP32);
  sz.add<int>(np1); sz.add<int>(np1); sz.add<float>(np1);
  sz.add<int>(G + 1); sz.add<int>(nnbr); sz.add<double>(G);
  const bool mfma = mcpg::env_group("MCPG_ADJ_MFMA", 1) != 0;
  if (mfma) sz.add<float>(G * bins * L.Sp);
  sz.add<float>(np1 * bins * bins); sz.add<double>(np1);
  sz.add<double>(G * bins); sz.add<float>(G * S); sz.add<float>(G); sz.add<int>(G); sz.add<int>(G);
  void* base = mcpg::workspace(h, sz.off, &st);
  if (!base) return st;
  mcpg::Carver cv(base);
  const float* dXin = X;
  int64_t lx = ldx;
  float* ddX = dX;
  int64_t ldd = lddx;
  if (ptrs == MCPG_HOST) {
    float* t = cv.take<float>(G * S);
    if ((st = mcpg::copy_2d(h, t, S * 4, X, ldx * 4, S * 4, G, hipMemcpyHostToDevice))) return st;
    dXin = t; lx = S;
    ddX = cv.take<float>(G * S); ldd = S;
  }
  WeightBufs wb = take_weight_bufs(cv, L, MCPG_FP32);
  // pair list, CSR and per-gene weight sums: one block, built on the host and uploaded at once
  int* d_pi = cv.take<int>(np1);
  int* d_pj = cv.take<int>(np1);
  float* d_m = cv.take<float>(np1);
  int* d_off = cv.take<int>(G + 1);
  int* d_nbr = cv.take<int>(nnbr);
  double* d_cw = cv.take<double>(G);
  char* const dblk = reinterpret_cast<char*>(d_pi);
  const size_t blk_bytes = (size_t)(reinterpret_cast<char*>(d_cw + G) - dblk);
  float* d_accg = mfma ? cv.take<float>(G * bins * L.Sp) : nullptr;
  float* d_Gm = cv.take<float>(np1 * bins * bins);
  double* d_Lp = cv.take<double>(np1);
  double* d_pm = cv.take<double>(G * bins);
  float* d_xs = cv.take<float>(G * S);
  float* d_scl = cv.take<float>(G);
  int* d_amin = cv.take<int>(G);
  int* d_amax = cv.take<int>(G);
  {
    // the block in the handle's pinned staging buffer when it fits (else in pageable memory)
    std::vector<char> pageable;
    char* hblk = static_cast<char*>(mcpg::pinned_staging(h, 0, blk_bytes, &st));
    if (st) return st;
    if (!hblk) {
      pageable.resize(blk_bytes);
      hblk = pageable.data();
    }
    auto host_of = [&](const void* d) { return hblk + (reinterpret_cast<const char*>(d) - dblk); };
    int* hnbr = reinterpret_cast<int*>(host_of(d_nbr));
    double* hcw = reinterpret_cast<double*>(host_of(d_cw));
    if (npairs > 0) {
      std::memcpy(host_of(d_pi), pi, npairs * 4);
      std::memcpy(host_of(d_pj), pj, npairs * 4);
      std::memcpy(host_of(d_m), m, npairs * 4);
    }
    std::memcpy(host_of(d_off), off.data(), (G + 1) * 4);
    std::vector<int> fill(off.begin(), off.end() - 1);
    for (int64_t n = 0; n < npairs; ++n) {
      int e = fill[pi[n]]++;
      hnbr[3 * e] = pj[n]; hnbr[3 * e + 1] = (int)n; hnbr[3 * e + 2] = 0;
      e = fill[pj[n]]++;
      hnbr[3 * e] = pi[n]; hnbr[3 * e + 1] = (int)n; hnbr[3 * e + 2] = 1;
    }
    hnbr[nnbr - 3] = hnbr[nnbr - 2] = hnbr[nnbr - 1] = 0;
    for (int64_t g = 0; g < G; ++g) {   // per gene: sum of m over its entries, in CSR order
      double c = 0.0;
      for (int e = off[g]; e < off[g + 1]; ++e) c += (double)m[hnbr[3 * e + 1]];
      hcw[g] = c;
    }
    if (pageable.empty()) {
      if ((st = mcpg::pinned_upload(h, 0, dblk, blk_bytes))) return st;
    } else {
      MCPG_HIP(h, hipMemcpy(dblk, hblk, blk_bytes, hipMemcpyHostToDevice));
    }
  }
  if ((st = clear_weight_bufs(h, wb, L))) return st;
  if ((st = mcpg::reset_flags(h))) return st;
  PrepOut o{};
  o.w32 = wb.w32; o.H64 = wb.H64; o.pmarg = d_pm; o.xs = d_xs; o.scl = d_scl; o.amin = d_amin; o.amax = d_amax;
  o.flags = h->d_flags;
  if ((st = prep_launch(h, dXin, lx, L, sp, 1, o))) return st;
  if (npairs > 0) {
    PairArgs pa;
    pa.w32 = wb.w32; pa.L = L; pa.pi = d_pi; pa.pj = d_pj; pa.m = d_m; pa.npairs = npairs; pa.H64 = wb.H64;
    pa.norm = 1.0f / (float)S; pa.Gm = d_Gm; pa.Lpart = d_Lp;
    mcpg::LaunchScope ls(h, mfma ? "bspline_pair_adjoint_mfma_kernel" : "bspline_pair_adjoint_kernel");
    if (mfma)
      hipLaunchKernelGGL(bspline_pair_adjoint_mfma_kernel, dim3((unsigned)ceil_div(npairs, 4)), dim3(256), 0, h->stream,
                         pa);
    else
      hipLaunchKernelGGL(bspline_pair_adjoint_kernel<16>, dim3((unsigned)ceil_div(npairs, 4)), dim3(256), 0, h->stream,
                         pa);
  }
  GeneArgs ga;
  ga.w32 = wb.w32; ga.L = L; ga.sp = sp; ga.off = d_off; ga.nbr = d_nbr; ga.cw = d_cw; ga.accg = d_accg; ga.Gm = d_Gm; ga.m = d_m; ga.pmarg = d_pm;
  ga.xs = d_xs; ga.scl = d_scl; ga.amin = d_amin; ga.amax = d_amax; ga.dX = ddX; ga.lddx = ldd;
  if (mfma) {
    {
      mcpg::LaunchScope ls(h, "bspline_gene_accumulate_kernel");
      hipLaunchKernelGGL(bspline_gene_accumulate_kernel, dim3((unsigned)ceil_div(L.Sp, kAdjPass), (unsigned)G),
                         dim3(256), 0, h->stream, ga);
    }
    mcpg::LaunchScope ls(h, "bspline_gene_epilogue_kernel");
    hipLaunchKernelGGL(bspline_gene_epilogue_kernel, dim3((unsigned)G), dim3(256), 0, h->stream, ga);
  } else {
    mcpg::LaunchScope ls(h, "bspline_gene_adjoint_kernel");
    hipLaunchKernelGGL((bspline_gene_adjoint_kernel<16, 4>), dim3((unsigned)G), dim3(256), 0, h->stream, ga);
  }
  if (ptrs == MCPG_HOST && (st = mcpg::copy_2d(h, dX, lddx * 4, ddX, S * 4, S * 4, G, hipMemcpyDeviceToHost))) return st;
  if (L_out) {
    // per-pair loss terms through the second pinned buffer when they fit, summed in pair order
    std::vector<double> pageable;
    double* lp = static_cast<double*>(mcpg::pinned_staging(h, 1, (size_t)np1 * 8, &st));
    if (st) return st;
    if (!lp) {
      pageable.assign(np1, 0.0);
      lp = pageable.data();
    }
    if (npairs > 0) MCPG_HIP(h, hipMemcpyAsync(lp, d_Lp, npairs * 8, hipMemcpyDeviceToHost, h->stream));
    if ((st = mcpg::finish(h, "mcpg_bspline_mi_backward_f32"))) return st;
    double Ls = 0.0;
    for (int64_t n = 0; n < npairs; ++n) Ls += lp[n];
    *L_out = Ls;
  }
  if ((st = mcpg::finish(h, "mcpg_bspline_mi_backward_f32"))) return st;
  return check_flags_nonfinite(h, "mcpg_bspline_mi_backward_f32");
}

}  // extern "C"
