// Adaptive-partitioning mutual information on rank-transformed rows, plus the rank, jitter and
// cumulative-histogram transforms that feed it.
//
// One wavefront evaluates one gene pair at a time, taking pairs in runs of 16 from a global
// counter.  Every window the reference's recursion can pop is a cell of one dyadic grid over the
// two rank ranges, so the wave first counts the pair's points on that grid to depth 5 and sums
// them into a quadrant-count pyramid in LDS.  The reference's chi-square split test is decided
// exactly in integers, and the LIFO walk over the windows above depth 5 is evaluated level by
// level with one lane per window.  Pairs whose walk goes deeper run it window by window: a LIFO
// stack in the wave's registers (entry k in lane k), the points bucketed by y rank in LDS
// (xGroupedByY), and each popped window's four quadrant counts from ballots and popcounts over
// its y range.  Counts are integers and match the reference exactly; leaf terms are evaluated a
// lane each, batched across pairs, and enter each pair's FP64 sum one by one in the reference's
// LIFO order.
#include <cmath>

#include "internal.h"

#pragma clang fp contract(off)

namespace {

using mcpg::ceil_div;

constexpr int kMaxWaves = 4;       // waves per workgroup (fewer when S is large)
constexpr int64_t kLdsBudget = 64 * 1024;
constexpr int64_t kMaxS = kLdsBudget / 8 - 1;  // keeps the per-gene histogram and bucket kernels in 32 KB of LDS
// Stack capacity = 64 * kStackSlots windows.  A pop pushes at most four children, so the stack
// grows by at most 3 per level, and every split halves both rank ranges: at most ceil(log2 S)
// = 13 levels for S <= 8191, so at most 40 entries (the overflow flag stays as a guard).
constexpr int kStackSlots = 1;
static_assert(kStackSlots == 1, "the stack's push and top address lane sp directly: one 64-lane slot");

struct Win {
  int x1, y1, x2, y2, n;
  int node = 0, depth = 0;   // the window's cell in the quadrant pyramid (node of level `depth`)
};

// Quadrant pyramid.  Every window the recursion can pop is a cell of a fixed dyadic grid: the
// root is [0, S) x [0, S) and each split halves both rank ranges at (lo + hi) >> 1, so a window
// at depth d is one of the 4^d cells whose x and y ranges are depth-d intervals of that bisection
// of [0, S).  A pair's point counts on the depth-kPyrDepth grid, summed up level by level, give
// the four quadrant counts of every window above that depth with one LDS read, exactly as the
// scans would count them.  Cells are numbered in Morton order (x bit below y bit at each level),
// so the four children of node k of level L are nodes 4k .. 4k + 3 of level L + 1, in the order
// of the reference's quadrants (x low / y low, x high / y low, x low / y high, x high / y high).
constexpr int kPyrDepth = 5;
__host__ __device__ constexpr int pyr_offset(int level) { return ((1 << (2 * level)) - 4) / 3; }   // level >= 1
constexpr int kPyrCells = pyr_offset(kPyrDepth + 1);   // levels 1 .. kPyrDepth: 4 + 16 + ... + 4^5 = 1364
// Counts are 16-bit (a window holds at most S <= 8191 points); level offsets are multiples of 4
// cells, so the four children of a node are one aligned 8-byte LDS read.

__device__ __forceinline__ int rfl(int v) { return __builtin_amdgcn_readfirstlane(v); }

// SplitMix64 (Steele, Lea, Flood 2014), counter-based: draw `ctr` of stream `seed`.
__host__ __device__ __forceinline__ uint64_t splitmix64(uint64_t seed, uint64_t ctr) {
  uint64_t z = seed + (ctr + 1) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// Sample s of gene g gets draw g * S + s of the stream; rows by blockIdx.y (gridDim.y <= 65535
// per launch, offset by g0), samples by blockIdx.x, so no 64-bit division per element.
__global__ void noise_kernel(const float* __restrict__ X, int64_t ldx, int64_t g0, int64_t S, uint64_t seed,
                             float* __restrict__ Y, int64_t ldy) {
  const int64_t g = g0 + blockIdx.y;
  for (int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; s < S; s += (int64_t)gridDim.x * blockDim.x) {
    const double u = (double)(splitmix64(seed, (uint64_t)(g * S + s)) >> 11) * 0x1.0p-53;
    const float j = (float)(u * 1e-5);
    Y[g * ldy + s] = X[g * ldx + s] + j;
  }
}

// Descending rank, ties by ascending index: r_s = #{t : x_t > x_s} + #{t < s : x_t == x_s}, i.e.
// the position of s when the samples are sorted by (value descending, index ascending).  One
// block per gene sorts the 64-bit keys (~u(x_s) << 32 | s) ascending with a bitonic network in
// LDS, where u is the unsigned order key of the float (-0 taken as +0, so they tie as under ==);
// n = S rounded up to a power of two, padding keys all ones.
__global__ __launch_bounds__(256) void rank_kernel(const float* __restrict__ X, int64_t ldx, int64_t S, int n,
                                                    int32_t* __restrict__ R, int64_t ldr, int* flags) {
  extern __shared__ uint64_t key[];
  const int64_t g = blockIdx.x;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    uint64_t k = ~0ull;
    if (i < S) {
      const float v = X[g * ldx + i];
      if (!isfinite(v)) atomicOr(flags, 1);
      const uint32_t b = __float_as_uint(v + 0.0f);   // -0 + 0 = +0
      const uint32_t u = b ^ ((b >> 31) ? 0xffffffffu : 0x80000000u);
      k = ((uint64_t)~u << 32) | (uint32_t)i;
    }
    key[i] = k;
  }
  __syncthreads();
  for (int k = 2; k <= n; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const int l = i ^ j;
        if (l > i) {
          const uint64_t a = key[i], b = key[l];
          if ((a > b) == ((i & k) == 0)) {
            key[i] = b;
            key[l] = a;
          }
        }
      }
      __syncthreads();
    }
  }
  for (int i = threadIdx.x; i < S; i += blockDim.x) R[g * ldr + (int64_t)(key[i] & 0xffffffffu)] = i;
}

// IntegralCumulativeHistogramKernel (AdaptivePartitioningMI.hpp:422-439) for one row per block.
// ident (nullable): ident[g] = 1 if the row's histogram is the identity (the ranks are a
// permutation of 0 .. S-1), else 0.
__global__ __launch_bounds__(256) void cumhist_kernel(const int32_t* __restrict__ R, int64_t ldr, int64_t S,
                                                       int32_t* __restrict__ H, int64_t ldh, int* flags,
                                                       int32_t* __restrict__ ident) {
  extern __shared__ int hsh[];
  const int64_t g = blockIdx.x;
  for (int64_t i = threadIdx.x; i <= S; i += blockDim.x) hsh[i] = 0;
  __syncthreads();
  for (int64_t s = threadIdx.x; s < S; s += blockDim.x) {
    const int r = R[g * ldr + s];
    if (r < 0 || r >= S) { atomicOr(flags, 2); continue; }
    atomicAdd(&hsh[r + 1], 1);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int64_t i = 2; i <= S; ++i) hsh[i] += hsh[i - 1];
  }
  __syncthreads();
  int same = 1;
  for (int64_t i = threadIdx.x; i <= S; i += blockDim.x) {
    H[g * ldh + i] = hsh[i];
    same &= hsh[i] == (int)i;
  }
  if (ident) {
    same = __syncthreads_and(same);
    if (threadIdx.x == 0) ident[g] = same;
  }
}

// The y-side of _createSparseMatrix (AdaptivePartitioningMI.hpp:462-493) once per gene instead of
// once per pair: P[g][cH[r] + t] = s for the t-th sample s of rank r, so that a pair's x ranks
// grouped by y rank are X[g_x][P[g_y][k]].  Within one rank the order is the atomics' (any order
// gives the same quadrant counts).
__global__ __launch_bounds__(256) void ybucket_kernel(const int32_t* __restrict__ R, int64_t ldr, int64_t S,
                                                      const int32_t* __restrict__ H, int64_t ldh,
                                                      int32_t* __restrict__ P) {
  extern __shared__ int cnt[];
  const int64_t g = blockIdx.x;
  for (int64_t i = threadIdx.x; i < S; i += blockDim.x) cnt[i] = 0;
  __syncthreads();
  for (int64_t s = threadIdx.x; s < S; s += blockDim.x) {
    const int r = R[g * ldr + s];
    P[g * S + H[g * ldh + r] + atomicAdd(&cnt[r], 1)] = (int32_t)s;
  }
}

// Z[g][s]: the bisection path of rank R[g][s] down kPyrDepth levels (bit kPyrDepth - 1 - l set
// when the rank falls in the upper half at level l + 1), its bits spread to the even positions
// of a Morton code.  A pair's cell at the deepest level is Z[gx][s] | Z[gy][s] << 1.
// Rows are padded to ldz (a multiple of the pyramid builder's 256-sample step) with code 0.
constexpr int kZStep = 256;
__host__ __device__ inline int64_t ap_ldz(int64_t S) { return (S + kZStep - 1) / kZStep * kZStep; }
__global__ void apcode_kernel(const int32_t* __restrict__ R, int64_t ldr, int64_t G, int64_t S,
                              uint16_t* __restrict__ Z, int64_t ldz) {
  const int64_t total = G * ldz;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t g = t / ldz, s = t % ldz;
    if (s >= S) {
      Z[t] = 0;
      continue;
    }
    const int r = R[g * ldr + s];
    int lo = 0, hi = (int)S, code = 0;
    for (int l = 0; l < kPyrDepth; ++l) {
      const int m = (lo + hi) >> 1;
      const int up = r >= m;
      code |= up << (2 * (kPyrDepth - 1 - l));
      if (up) lo = m;
      else hi = m;
    }
    Z[t] = (uint16_t)code;
  }
}

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
}

// The pair's point counts on the deepest grid (LDS atomics on the 32-bit words that hold two
// 16-bit cells), then each coarser level as sums of four children.
__device__ __forceinline__ uint32_t sum4_u16(uint2 q) {
  return (q.x & 0xffffu) + (q.x >> 16) + (q.y & 0xffffu) + (q.y >> 16);
}
__device__ __forceinline__ void build_pyramid(uint16_t* __restrict__ pyr, const uint16_t* __restrict__ zx,
                                              const uint16_t* __restrict__ zy, int S, int lane) {
  static_assert(pyr_offset(kPyrDepth) % 2 == 0, "the deepest level starts on a 32-bit word");
  uint32_t* leaf = reinterpret_cast<uint32_t*>(pyr + pyr_offset(kPyrDepth));
#pragma unroll
  for (int i = lane; i < (1 << (2 * kPyrDepth)) / 2; i += 64) leaf[i] = 0u;
  wave_sync();
  // the rows of Z are padded to a multiple of kZStep samples with code 0: every lane adds its
  // four points without a bounds test, and cell 0 gives the padding back afterwards
  const int Sp = (int)ap_ldz(S);
  for (int base = 0; base < Sp; base += kZStep) {
    uint32_t zxv[4], zyv[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      zxv[u] = zx[base + 64 * u + lane];
      zyv[u] = zy[base + 64 * u + lane];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) atomicAdd(&leaf[(zxv[u] >> 1) | zyv[u]], 1u << ((zxv[u] & 1u) << 4));
  }
  wave_sync();
  if (lane == 0 && Sp > S) leaf[0] -= (uint32_t)(Sp - S);
  wave_sync();
#pragma unroll
  for (int L = kPyrDepth - 1; L >= 1; --L) {
    const uint2* dn = reinterpret_cast<const uint2*>(pyr + pyr_offset(L + 1));
    uint16_t* up = pyr + pyr_offset(L);
    for (int k = lane; k < (1 << (2 * L)); k += 64) up[k] = (uint16_t)sum4_u16(dn[k]);
    wave_sync();
  }
}

__device__ __forceinline__ void tri_decode(int64_t Lid, int64_t nb, int64_t* bm, int64_t* bn) {
  double a = 2.0 * (double)nb + 1.0;
  int64_t r = (int64_t)((a - sqrt(a * a - 8.0 * (double)Lid)) * 0.5);
  if (r < 0) r = 0;
  if (r > nb - 1) r = nb - 1;
  auto start = [nb](int64_t rr) { return rr * nb - rr * (rr - 1) / 2; };
  while (r > 0 && start(r) > Lid) --r;
  while (r + 1 < nb && start(r + 1) <= Lid) ++r;
  *bm = r;
  *bn = r + (Lid - start(r));
}

struct APArgs {
  const int32_t* R;
  int64_t ldr;
  const int32_t* H;   // cumulative histograms, G x (S+1)
  int64_t ldh;
  const int32_t* P;   // samples grouped by rank, G x S (ybucket_kernel)
  int64_t G, S;
  float* MI;
  double* MI64;       // non-null: the MI kept in double (AdaptivePartitionRankMIKernel2<IT, double>)
  int64_t ldmi;
  // output block: MI holds rows r0 .. r0 + nr - 1 and columns c0 .. c0 + nc - 1 of the G x G
  // matrix; block = false: the whole matrix, its pairs enumerated over the upper triangle
  int64_t r0, nr, c0, nc;
  bool block;
  // debug / cells mode
  const int32_t* px;
  const int32_t* py;
  int64_t npairs;     // > 0: cells mode over the listed pairs
  int32_t max_leaves;
  int32_t* cells;
  int32_t* nleaves;
  double* mi_out;
  int* flags;
  unsigned long long* queue;   // all-pairs mode: next unclaimed pair index, zero at launch
  const int32_t* ident;        // per gene: 1 if its histogram is the identity (cumhist_kernel)
  const uint16_t* Z;           // per gene and sample: the Morton x code of its rank's pyramid cell (apcode_kernel)
  int64_t ldz;
};

constexpr int kChunk = 16;     // consecutive pairs a wave claims from the queue at a time

// Quadrant counts of window w split at (mx, my) over xg[b0 .. b1), lower/upper y halves split at bm.
// The lower half [b0, bm) and the upper half [bm, b1) are scanned separately, 64 points at a
// time: two ballots and two scalar popcounts per 64 points.  x in [lo, hi) is tested as
// (unsigned)(x - lo) < hi - lo, one compare per ballot; lanes past the half read the array's
// padding (kXgPad entries past S) and take x = -1, which no window contains.  (One pass with
// four ballots per 64 points measured slower: 9.5 against 8.3 ms at 2000 x 1000.)
constexpr int kXgPad = 64;
// LDS bytes per wave: the quadrant pyramid, then the grouped points (16-bit ranks), each rounded
// to 16 bytes
__host__ __device__ constexpr int node_offset(int level) { return ((1 << (2 * level)) - 1) / 3; }
constexpr int kTreeNodes = node_offset(kPyrDepth);   // levels 0 .. kPyrDepth - 1: 341 nodes
constexpr int kStageItems = 64;
__host__ __device__ inline int64_t ap_pyr_bytes() { return (2 * kPyrCells + 15) & ~15; }
// the tree pass's staged leaves and node words share the space of the grouped points, which
// only the sequential walk uses
__host__ __device__ inline int64_t ap_tree_bytes() { return (8 * kStageItems + 4 * kTreeNodes + 15) & ~15; }
__host__ __device__ inline int64_t ap_wave_bytes(int64_t S) {
  const int64_t xg = (2 * (S + kXgPad) + 15) & ~(int64_t)15;
  return ap_pyr_bytes() + (xg > ap_tree_bytes() ? xg : ap_tree_bytes());
}

__device__ __forceinline__ void count_half(const uint16_t* __restrict__ xg, int b0, int b1, int x1, int mx, int x2, int lane,
                                           int& nl, int& nr) {
  nl = nr = 0;
  const uint32_t wl = (uint32_t)(mx - x1), wr = (uint32_t)(x2 - mx);
  for (int base = b0; base < b1; base += 64) {
    const int idx = base + lane;
    const int v = (int)xg[idx];
    const int x = idx < b1 ? v : -1;
    nl += __popcll(__builtin_amdgcn_ballot_w64((uint32_t)(x - x1) < wl));
    nr += __popcll(__builtin_amdgcn_ballot_w64((uint32_t)(x - mx) < wr));
  }
}
__device__ __forceinline__ void count_quadrants(const uint16_t* __restrict__ xg, int b0, int bm, int b1, int x1, int mx,
                                                int x2, int lane, int& n0, int& n1, int& n2, int& n3) {
  count_half(xg, b0, bm, x1, mx, x2, lane, n0, n1);
  count_half(xg, bm, b1, x1, mx, x2, lane, n2, n3);
}

// The reference's test (AdaptivePartitioningMI.hpp:608-619) computes, in double, e = n / 4,
// chi = sum (n_i - e)^2 / e and splits when chi > 7.815.  Every step before the division is exact,
// so chi is D / (4 n) rounded once, with D = sum (4 n_i - n)^2 an integer (D <= 12 n^2 < 2^31).
// A correctly rounded quotient exceeds the double c = 7.815 = M 2^-50 exactly when the exact one
// exceeds the midpoint (2M + 1) 2^-51 between c and the next double (it never equals it: that
// would need 2^49 | n), i.e. when D > floor(n (2M + 1) / 2^49).
constexpr double kChi2 = 7.815;
constexpr uint64_t kChi2Mid = 2 * (uint64_t)(kChi2 * 0x1p50) + 1;   // 2M + 1: 54 bits
static_assert((double)(kChi2Mid >> 1) * 0x1p-50 == kChi2, "7.815 is M 2^-50 for an integer M");
__host__ __device__ __forceinline__ int chi2_threshold(int n) {
  const uint64_t hi = (uint64_t)n * (kChi2Mid >> 32), lo = (uint64_t)n * (kChi2Mid & 0xffffffffull);
  return (int)((hi + (lo >> 32)) >> 17);
}

__device__ __forceinline__ double readlane_f64(double v, int k) {
  const uint64_t b = __builtin_bit_cast(uint64_t, v);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)b, k);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(b >> 32), k);
  return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

// Rank intervals of node k of pyramid level L: the bisection of [0, S) along the node's Morton
// path (x bit below y bit at each level).
__device__ __forceinline__ void node_bounds(int S, int L, int k, int& x1, int& x2, int& y1, int& y2) {
  int xl = 0, xh = S, yl = 0, yh = S;
  for (int l = 0; l < L; ++l) {
    const int sh = 2 * (L - 1 - l);
    const int mx = (xl + xh) >> 1, my = (yl + yh) >> 1;
    if ((k >> sh) & 1) xl = mx; else xh = mx;
    if ((k >> (sh + 1)) & 1) yl = my; else yh = my;
  }
  x1 = xl; x2 = xh; y1 = yl; y2 = yh;
}

// The LIFO walk of the reference over the windows above the pyramid's deepest level, evaluated
// level by level with one lane per node instead of one window at a time.  A node's word in LDS:
// bit 0 visited (the walk pops it), bit 1 split, bits 2-5 the children it pushes, bits 6-9 the
// children it adds as leaves at once, bits 16-23 the number of leaves its subtree adds, bits
// 24-31 the position of its first leaf in the walk's order.  Popping a split window adds its
// non-splittable children with points (children 0 .. 3 in order), then walks its pushed
// children last-pushed first, so a split node's leaves are [its leaf children] followed by the
// subtrees of its pushed children 3, 2, 1, 0.  The leaves (n, nx, ny) are staged in that order.
// Returns their number, or -1 when the walk would pop a window below the pyramid or add more than
// kStageItems leaves (the caller then walks the pair window by window).  Needs S >= 2^(kPyrDepth+1),
// so that every interval down to the pyramid's deepest level holds two ranks or more and a window
// is splittable exactly when it holds more than two points.
template <typename HX, typename HY>
__device__ int tree_pass(const uint16_t* __restrict__ pyr, uint32_t* __restrict__ nodes, uint2* __restrict__ stage,
                         int S, int lane, HX&& hx, HY&& hy) {
  auto child_counts = [&](int L, int k, int (&cc)[4]) {
    const uint2 c = *reinterpret_cast<const uint2*>(pyr + pyr_offset(L + 1) + 4 * k);
    cc[0] = (int)(c.x & 0xffffu); cc[1] = (int)(c.x >> 16); cc[2] = (int)(c.y & 0xffffu); cc[3] = (int)(c.y >> 16);
  };
  // 1. top-down: visited, split, pushed and leaf children
  int deepest = 0;
  for (int L = 0; L < kPyrDepth; ++L) {
    const int level_nodes = 1 << (2 * L);
    bool pushes = false;
    for (int k0 = 0; k0 < level_nodes; k0 += 64) {
      const int k = k0 + lane;
      if (k < level_nodes) {
        uint32_t word = 0;
        const bool visited = L == 0 || ((nodes[node_offset(L - 1) + (k >> 2)] >> (2 + (k & 3))) & 1u);
        if (visited) {
          const int n = L == 0 ? S : (int)pyr[pyr_offset(L) + k];
          int cc[4];
          child_counts(L, k, cc);
          int dev = 0;
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const int t = 4 * cc[i] - n;
            dev += t * t;
          }
          const bool split = L == 0 || dev > chi2_threshold(n);
          uint32_t pushed = 0, leafm = 0;
          if (split) {
#pragma unroll
            for (int i = 0; i < 4; ++i) {
              if (cc[i] > 2) pushed |= 1u << i;
              else if (cc[i] > 0) leafm |= 1u << i;
            }
          }
          word = 1u | ((uint32_t)split << 1) | (pushed << 2) | (leafm << 6);
          pushes |= pushed != 0;
        }
        nodes[node_offset(L) + k] = word;
      }
    }
    wave_sync();
    deepest = L;
    if (__builtin_amdgcn_ballot_w64(pushes) == 0) break;
    if (L == kPyrDepth - 1) return -1;   // a window below the pyramid would be popped
  }
  // 2. bottom-up: leaves per visited subtree (saturating at 255)
  for (int L = deepest; L >= 0; --L) {
    const int level_nodes = 1 << (2 * L);
    for (int k0 = 0; k0 < level_nodes; k0 += 64) {
      const int k = k0 + lane;
      if (k < level_nodes) {
        const uint32_t word = nodes[node_offset(L) + k];
        if (word & 1u) {
          int cnt = 1;
          if (word & 2u) {
            cnt = __builtin_popcount((word >> 6) & 0xfu);
            const uint32_t pushed = (word >> 2) & 0xfu;
#pragma unroll
            for (int i = 0; i < 4; ++i)
              if (pushed & (1u << i)) cnt += (int)((nodes[node_offset(L + 1) + 4 * k + i] >> 16) & 0xffu);
          }
          nodes[node_offset(L) + k] = word | ((uint32_t)(cnt < 255 ? cnt : 255) << 16);
        }
      }
    }
    wave_sync();
  }
  const int total = (int)((nodes[0] >> 16) & 0xffu);
  if (total > kStageItems) return -1;
  // 3. top-down: leaf positions; split nodes stage their leaf children and place their pushed ones
  for (int L = 0; L <= deepest; ++L) {
    const int level_nodes = 1 << (2 * L);
    for (int k0 = 0; k0 < level_nodes; k0 += 64) {
      const int k = k0 + lane;
      if (k < level_nodes) {
        const uint32_t word = nodes[node_offset(L) + k];
        if (word & 1u) {
          const int pos = (int)(word >> 24);
          int x1, x2, y1, y2;
          if (word & 2u) {
            int cc[4];
            child_counts(L, k, cc);
            const uint32_t leafm = (word >> 6) & 0xfu, pushed = (word >> 2) & 0xfu;
            int q = pos;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
              if (leafm & (1u << i)) {
                node_bounds(S, L + 1, 4 * k + i, x1, x2, y1, y2);
                stage[q++] = make_uint2((uint32_t)cc[i] | ((uint32_t)(hx(x2) - hx(x1)) << 16),
                                        (uint32_t)(hy(y2) - hy(y1)));
              }
            }
#pragma unroll
            for (int i = 3; i >= 0; --i) {
              if (pushed & (1u << i)) {
                uint32_t* cw = nodes + node_offset(L + 1) + 4 * k + i;
                const uint32_t c = *cw;
                *cw = (c & 0x00ffffffu) | ((uint32_t)q << 24);
                q += (int)((c >> 16) & 0xffu);
              }
            }
          } else {
            node_bounds(S, L, k, x1, x2, y1, y2);
            const int n = L == 0 ? S : (int)pyr[pyr_offset(L) + k];
            stage[pos] = make_uint2((uint32_t)n | ((uint32_t)(hx(x2) - hx(x1)) << 16), (uint32_t)(hy(y2) - hy(y1)));
          }
        }
      }
    }
    wave_sync();
  }
  return total;
}

// Leaf terms (compute_for_window, :525-538) of consecutive pairs are batched per wave: leaf k of
// the batch is held by lane k, and a batch (flushed before it can pass 64 leaves, when 64 pairs
// have completed, and at the end) evaluates n log2(n / (nx ny)) on all lanes at once.  Each
// pair's terms then enter its own FP64 sum one by one in leaf order, as in the reference's loop:
// completed pair k of the batch (lane k) keeps its first leaf, its sum over leaves of earlier
// batches and where its result goes.  One evaluation then serves several pairs.
struct LeafBatch {
  int pend = 0, pn = 0, pnx = 0, pny = 0;   // leaves; lane k: leaf k
  int done = 0;                              // completed pairs waiting for the evaluation
  int p_first = 0, p_x = 0, p_y = 0;         // lane k: pair k's first leaf and its genes (or output slot)
  double p_mi = 0.0;                         // lane k: pair k's sum over leaves of earlier batches
  int cur_first = 0;                         // first leaf of the pair in progress
  double cur_mi = 0.0;                       // its sum over leaves of earlier batches
};

// Evaluates the batch: the completed pairs' sums are finished and handed to store(x, y, value)
// (value = mi/S + log2S), the pair in progress keeps its partial sum.
template <typename Store>
__device__ __forceinline__ void flush_batch(LeafBatch& b, int lane, int S, double log2S, Store&& store) {
  double t = 0.0;
  if (lane < b.pend) t = (double)b.pn * log2((double)b.pn / (double)((int64_t)b.pnx * (int64_t)b.pny));
  int l = 0;
  for (int k = 0; k < b.done; ++k) {
    const int end = k + 1 < b.done ? __builtin_amdgcn_readlane(b.p_first, k + 1) : b.cur_first;
    double m = readlane_f64(b.p_mi, k);
    for (; l < end; ++l) m += readlane_f64(t, l);
    store(__builtin_amdgcn_readlane(b.p_x, k), __builtin_amdgcn_readlane(b.p_y, k), m / (double)S + log2S);
  }
  for (; l < b.pend; ++l) b.cur_mi += readlane_f64(t, l);
  b.pend = 0;
  b.done = 0;
  b.cur_first = 0;
}

// Evaluates AdaptivePartitionRankMIKernel2::operator()(x=gx, y=gy) (AdaptivePartitioningMI.hpp:553-645)
// into the wave's leaf batch; the pair is recorded there as completed under (ox, oy), and its value
// mi/S + log2S reaches `store` when the batch is evaluated.  If rec != nullptr, leaf (nxy, nx, ny)
// triples are recorded in summation order (up to max_leaves) and *nleaves receives the count.
template <typename Store>
__device__ void ap_pair(const APArgs& a, const int32_t* __restrict__ H, int64_t gx, int64_t gy, uint16_t* pyr,
                        uint16_t* xg, int lane, double log2S, LeafBatch& b, int ox, int oy, Store&& store, int32_t* rec,
                        int32_t* nleaves) {
  uint2* stage = reinterpret_cast<uint2*>(xg);
  uint32_t* nodes = reinterpret_cast<uint32_t*>(stage + kStageItems);
  const int S = (int)a.S;
  const int32_t* __restrict__ cHX = H + gx * a.ldh;
  const int32_t* __restrict__ cHY = H + gy * a.ldh;
  const int32_t* __restrict__ xr = a.R + gx * a.ldr;
  // a gene whose ranks are a permutation has the identity histogram: its entries need no load
  const bool idX = rfl(a.ident[gx]) != 0, idY = rfl(a.ident[gy]) != 0;
  auto hx = [&](int i) { return idX ? i : cHX[i]; };
  auto hy = [&](int i) { return idY ? i : cHY[i]; };

  build_pyramid(pyr, a.Z + gx * a.ldz, a.Z + gy * a.ldz, S, lane);

  auto complete = [&]() {   // the pair is complete: record it in the batch
    if (lane == b.done) {
      b.p_first = b.cur_first;
      b.p_mi = b.cur_mi;
      b.p_x = ox;
      b.p_y = oy;
    }
    ++b.done;
    b.cur_first = b.pend;
    b.cur_mi = 0.0;
    if (b.done == 64) flush_batch(b, lane, S, log2S, store);
  };
  if (S >= (2 << kPyrDepth)) {
    const int nl = tree_pass(pyr, nodes, stage, S, lane, hx, hy);
    if (nl >= 0) {
      if (b.pend + nl > 64) flush_batch(b, lane, S, log2S, store);
      if (lane < nl) {
        const uint2 it = stage[lane];
        if (rec && lane < a.max_leaves) {
          rec[3 * lane + 0] = (int32_t)(it.x & 0xffffu);
          rec[3 * lane + 1] = (int32_t)(it.x >> 16);
          rec[3 * lane + 2] = (int32_t)it.y;
        }
      }
      const int src = lane - b.pend;
      if (src >= 0 && src < nl) {
        const uint2 it = stage[src];
        b.pn = (int)(it.x & 0xffffu);
        b.pnx = (int)(it.x >> 16);
        b.pny = (int)it.y;
      }
      b.pend += nl;
      if (nleaves) *nleaves = nl;
      wave_sync();   // the staged leaves are read before the next pair reuses the space
      complete();
      return;
    }
  }

  // _createSparseMatrix (:462-493): x ranks grouped by y rank, through the y gene's bucket order
  // (P holds sample indices in [0, S)); eight loads per lane in flight at a time.  Only windows
  // below the pyramid scan points, so a pair builds this on its first such window.
  bool grouped = false;
  auto group_points = [&]() {
    const int32_t* __restrict__ py = a.P + gy * S;
    for (int base = 0; base < S; base += 64 * 8) {
      int v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int i = base + 64 * u + lane;
        v[u] = i < S ? py[i] : 0;
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int i = base + 64 * u + lane;
        v[u] = i < S ? xr[v[u]] : 0;
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int i = base + 64 * u + lane;
        if (i < S) xg[i] = (uint16_t)v[u];
      }
    }
    wave_sync();
    grouped = true;
  };

  // register-resident LIFO stack: entry k in lane k, with x1, x2 and y1, y2 packed as 16-bit
  // halves (ranks are below 2^13) and n, the pyramid node and the depth in the third register
  // (13, 10 and 5 bits): a push writes three registers in one lane, a read is three readlanes
  int sxs = 0, sys = 0, sn = 0;
  int sp = 0;  // number of entries
  bool whole = true;
  auto push = [&](const Win& w) {
    if (lane == sp) {
      sxs = (w.x1 << 16) | w.x2;
      sys = (w.y1 << 16) | w.y2;
      sn = w.n | (w.node << 13) | (w.depth << 23);
    }
    ++sp;
  };
  auto top = [&]() {
    const int ln = sp - 1;
    const int xs = __builtin_amdgcn_readlane(sxs, ln), ys = __builtin_amdgcn_readlane(sys, ln);
    const int nn = __builtin_amdgcn_readlane(sn, ln);
    Win w{xs >> 16, ys >> 16, xs & 0xffff, ys & 0xffff, nn & 0x1fff};
    w.node = (nn >> 13) & 0x3ff;
    w.depth = nn >> 23;
    return w;
  };
  int32_t leaf = 0;
  auto term = [&](const Win& w) {
    const int nx = hx(w.x2) - hx(w.x1);
    const int ny = hy(w.y2) - hy(w.y1);
    if (rec && lane == 0 && leaf < a.max_leaves) {
      rec[3 * leaf + 0] = w.n;
      rec[3 * leaf + 1] = nx;
      rec[3 * leaf + 2] = ny;
    }
    ++leaf;
    if (lane == b.pend) { b.pn = w.n; b.pnx = nx; b.pny = ny; }
    ++b.pend;
  };

  push(Win{0, 0, S, S, S});
  int overflow = 0;
  while (sp > 0) {
    if (b.pend > 60) flush_batch(b, lane, S, log2S, store);   // a window adds at most four leaves
    const Win w = top();
    const int mx = (w.x1 + w.x2) >> 1;
    const int my = (w.y1 + w.y2) >> 1;
    Win q[4] = {Win{w.x1, w.y1, mx, my, 0}, Win{mx, w.y1, w.x2, my, 0}, Win{w.x1, my, mx, w.y2, 0},
                Win{mx, my, w.x2, w.y2, 0}};
    if (w.depth < kPyrDepth) {
      // the window's four quadrants are the children of its pyramid node
      const uint2 c = *reinterpret_cast<const uint2*>(pyr + pyr_offset(w.depth + 1) + 4 * w.node);
      const uint32_t c0 = (uint32_t)rfl((int)c.x), c1 = (uint32_t)rfl((int)c.y);
      q[0].n = (int)(c0 & 0xffffu); q[1].n = (int)(c0 >> 16); q[2].n = (int)(c1 & 0xffffu); q[3].n = (int)(c1 >> 16);
    } else {
      if (!grouped) group_points();
      const int b0 = rfl(hy(w.y1)), bm = rfl(hy(my)), b1 = rfl(hy(w.y2));
      count_quadrants(xg, b0, bm, b1, w.x1, mx, w.x2, lane, q[0].n, q[1].n, q[2].n, q[3].n);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      q[i].depth = w.depth + 1;
      q[i].node = w.depth + 1 <= kPyrDepth ? 4 * w.node + i : 0;
    }
    // chi-square > 7.815 (:608-619), decided exactly in integer arithmetic
    int dev = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int t = 4 * q[i].n - w.n;
      dev += t * t;
    }
    if (dev > chi2_threshold(w.n) || whole) {
      --sp;
      whole = false;  // only the first window is `whole`; pushed children are copies with whole = false
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const Win& c = q[i];
        const bool splittable = ((c.x1 + 1 < c.x2) || (c.y1 + 1 < c.y2)) && (c.n > 2);
        if (splittable) {
          if (sp >= 64) { overflow = 1; continue; }
          push(c);
        } else if (c.n > 0) {
          term(c);
        }
      }
    } else if (w.n > 0) {
      term(w);
      --sp;
    } else {
      --sp;  // unreachable for pushed windows (n > 2); keeps the loop finite
    }
  }
  if (overflow && lane == 0) atomicOr(a.flags, 4);
  if (nleaves) *nleaves = leaf;
  complete();
}

// H is passed on its own, restrict-qualified: the histogram entries a window reads are
// wave-uniform and, with no store that may alias them, load through the scalar cache.  CELLS:
// the listed pairs with their leaves recorded; otherwise every pair of the upper triangle.
template <bool CELLS>
__global__ __launch_bounds__(64 * kMaxWaves) __attribute__((amdgpu_waves_per_eu(8)))
void ap_mi_kernel(APArgs a, const int32_t* __restrict__ H) {
  extern __shared__ int lds[];
  const int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int waves = blockDim.x >> 6;
  uint16_t* pyr = reinterpret_cast<uint16_t*>(reinterpret_cast<char*>(lds) + (int64_t)w * ap_wave_bytes(a.S));
  uint16_t* xg = reinterpret_cast<uint16_t*>(reinterpret_cast<char*>(pyr) + ap_pyr_bytes());
  const int64_t total_waves = (int64_t)gridDim.x * waves;
  const int64_t wave_id = (int64_t)blockIdx.x * waves + w;
  const double log2S = log2((double)a.S);
  const int S = (int)a.S;
  LeafBatch b;
  if constexpr (CELLS) {
    auto store = [&](int n, int, double v) {
      if (lane == 0) a.mi_out[n] = v;
    };
    for (int64_t n = wave_id; n < a.npairs; n += total_waves) {
      int32_t nl = 0;
      ap_pair(a, H, a.px[n], a.py[n], pyr, xg, lane, log2S, b, (int)n, 0, store,
              a.cells + n * 3 * (int64_t)a.max_leaves, &nl);
      if (lane == 0) a.nleaves[n] = nl;
    }
    flush_batch(b, lane, S, log2S, store);
    return;
  } else {
    // pair (i, j), i <= j, goes to (i, j) and (j, i) where those fall inside the output block
    auto store = [&](int i, int j, double v) {
      if (lane == 0) {
        const float f = (float)v;
        const uint64_t ri = (uint64_t)(i - a.r0), cj = (uint64_t)(j - a.c0);
        const uint64_t rj = (uint64_t)(j - a.r0), ci = (uint64_t)(i - a.c0);
        const bool at_ij = ri < (uint64_t)a.nr && cj < (uint64_t)a.nc;
        const bool at_ji = i != j && rj < (uint64_t)a.nr && ci < (uint64_t)a.nc;
        if (a.MI64) {
          if (at_ij) a.MI64[(int64_t)ri * a.ldmi + (int64_t)cj] = v;
          if (at_ji) a.MI64[(int64_t)rj * a.ldmi + (int64_t)ci] = v;
        } else {
          if (at_ij) a.MI[(int64_t)ri * a.ldmi + (int64_t)cj] = f;
          if (at_ji) a.MI[(int64_t)rj * a.ldmi + (int64_t)ci] = f;
        }
      }
    };
    // All pairs (i <= j, row-major over the upper triangle; for a block, its positions row-major),
    // claimed kChunk at a time from a global counter so that waves whose pairs recurse less take
    // more of them; a wave leaves once the counter has passed the last pair.
    const int64_t P = a.block ? a.nr * a.nc : a.G * (a.G + 1) / 2;
    for (;;) {
      unsigned long long claim = 0;
      if (lane == 0) claim = atomicAdd(a.queue, (unsigned long long)kChunk);
      const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)claim);
      const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(claim >> 32));
      const int64_t base = (int64_t)(((uint64_t)hi << 32) | lo);
      if (base >= P) break;
      const int64_t end = base + kChunk < P ? base + kChunk : P;
      if (a.block) {
        // position (i, j) computes pair (min, max), the orientation of the whole-matrix launch;
        // a position below the diagonal whose mirror is also in the block is left to the mirror
        for (int64_t p = base; p < end; ++p) {
          const int64_t i = a.r0 + p / a.nc, j = a.c0 + p % a.nc;
          if (i > j && (uint64_t)(j - a.r0) < (uint64_t)a.nr && (uint64_t)(i - a.c0) < (uint64_t)a.nc) continue;
          const int64_t x = i < j ? i : j, y = i < j ? j : i;
          ap_pair(a, H, x, y, pyr, xg, lane, log2S, b, (int)x, (int)y, store, nullptr, nullptr);
        }
      } else {
        int64_t i, j;
        tri_decode(base, a.G, &i, &j);
        for (int64_t p = base; p < end; ++p) {
          ap_pair(a, H, i, j, pyr, xg, lane, log2S, b, (int)i, (int)j, store, nullptr, nullptr);
          if (++j == a.G) { ++i; j = i; }
        }
      }
    }
    flush_batch(b, lane, S, log2S, store);
  }
}

int check_flags(mcpg_ctx* h, const char* what) {
  int flags[1];
  int st = mcpg::read_flags(h, flags, 1);
  if (st) return st;
  if (flags[0] & 1) return mcpg::set_error(h, MCPG_ERR_NONFINITE, std::string(what) + ": non-finite input value");
  if (flags[0] & 2) return mcpg::set_error(h, MCPG_ERR_INVALID_ARG, std::string(what) + ": rank outside [0, S)");
  if (flags[0] & 4) return mcpg::set_error(h, MCPG_ERR_UNSUPPORTED, std::string(what) + ": window stack overflow");
  return MCPG_OK;
}

}  // namespace

extern "C" {

int mcpg_add_noise_f32(mcpg_handle h, int ptrs, const float* X, int64_t G, int64_t S, int64_t ldx, uint64_t seed,
                       float* Y, int64_t ldy) {
  MCPG_CHECK_HANDLE(h);
  mcpg::begin_call(h);
  if (G < 0 || S < 0) return mcpg::set_error(h, MCPG_ERR_SHAPE, "negative dimension");
  if (G == 0 || S == 0) return MCPG_OK;
  if (!X || !Y) return mcpg::set_error(h, MCPG_ERR_INVALID_ARG, "null matrix pointer");
  if (ldx < S || ldy < S) return mcpg::set_error(h, MCPG_ERR_INVALID_ARG, "leading dimension too small");
  if (ptrs != MCPG_HOST && ptrs != MCPG_DEVICE) return mcpg::set_error(h, MCPG_ERR_INVALID_ARG, "bad ptrs kind");
  MCPG_HIP(h, hipSetDevice(h->device));
  int st;
  const float* dX = X;
  float* dY = Y;
  int64_t lx = ldx, ly = ldy;
  if (ptrs == MCPG_HOST) {
    mcpg::Sizer sz;
    sz.add<float>(G * S);
    sz.add<float>(G * S);
    void* base = mcpg::workspace(h, sz.off, &st);
    if (!base) return st;
    mcpg::Carver cv(base);
    float* a = cv.take<float>(G * S);
    float* b = cv.take<float>(G * S);
    if ((st = mcpg::copy_2d(h, a, S * 4, X, ldx * 4, S * 4, G, hipMemcpyHostToDevice))) return st;
    dX = a; dY = b; lx = ly = S;
  }
  {
    mcpg::LaunchScope ls(h, "noise_kernel");
    const unsigned gx = (unsigned)std::min<int64_t>(ceil_div(S, 256), 1024);
    for (int64_t g0 = 0; g0 < G; g0 += 65535)
      hipLaunchKernelGGL(noise_kernel, dim3(gx, (unsigned)std::min<int64_t>(G - g0, 65535)), dim3(256), 0, h->stream,
                         dX, lx, g0, S, seed, dY, ly);
  }
  if (ptrs == MCPG_HOST && (st = mcpg::copy_2d(h, Y, ldy * 4, dY, S * 4, S * 4, G, hipMemcpyDeviceToHost))) return st;
  return mcpg::finish(h, "mcpg_add_noise_f32");
}

int mcpg_rank_f32(mcpg_handle h, int ptrs, const float* X, int64_t G, int64_t S, int64_t ldx, int32_t* R, int64_t ldr) {
  MCPG_CHECK_HANDLE(h);
  mcpg::begin_call(h);
  if (G < 0 || S < 0) return mcpg::set_error(h, MCPG_ERR_SHAPE, "negative dimension");
  if (G == 0 || S == 0) return MCPG_OK;
  if (!X || !R) return mcpg::set_error(h, MCPG_ERR_INVALID_ARG, "null matrix pointer");
  if (ldx < S || ldr < S) return mcpg::set_error(h, MCPG_ERR_INVALID_ARG, "leading dimension too small");
  if (S > 16384) return mcpg::set_error(h, MCPG_ERR_UNSUPPORTED, "rank supports S <= 16384");
  if (ptrs != MCPG_HOST && ptrs != MCPG_DEVICE) return mcpg::set_error(h, MCPG_ERR_INVALID_ARG, "bad ptrs kind");
  MCPG_HIP(h, hipSetDevice(h->device));
  int st;
  const float* dX = X;
  int32_t* dR = R;
  int64_t lx = ldx, lr = ldr;
  if (ptrs == MCPG_HOST) {
    mcpg::Sizer sz;
    sz.add<float>(G * S);
    sz.add<int32_t>(G * S);
    void* base = mcpg::workspace(h, sz.off, &st);
    if (!base) return st;
    mcpg::Carver cv(base);
    float* a = cv.take<float>(G * S);
    dR = cv.take<int32_t>(G * S);
    if ((st = mcpg::copy_2d(h, a, S * 4, X, ldx * 4, S * 4, G, hipMemcpyHostToDevice))) return st;
    dX = a; lx = lr = S;
  }
  if ((st = mcpg::reset_flags(h))) return st;
  {
    int n = 2;
    while (n < S) n <<= 1;
    static bool attr = false;   // up to 16384 keys of 8 bytes: 128 KB of LDS
    if (!attr) {
      MCPG_HIP(h, hipFuncSetAttribute(reinterpret_cast<const void*>(&rank_kernel),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, 16384 * 8));
      attr = true;
    }
    mcpg::LaunchScope ls(h, "rank_kernel");
    hipLaunchKernelGGL(rank_kernel, dim3((unsigned)G), dim3(256), (size_t)n * sizeof(uint64_t), h->stream, dX, lx, S,
                       n, dR, lr, h->d_flags);
  }
  if (ptrs == MCPG_HOST && (st = mcpg::copy_2d(h, R, ldr * 4, dR, S * 4, S * 4, G, hipMemcpyDeviceToHost))) return st;
  if ((st = mcpg::finish(h, "mcpg_rank_f32"))) return st;
  return check_flags(h, "mcpg_rank_f32");
}

int mcpg_cumhist_i32(mcpg_handle h, int ptrs, const int32_t* R, int64_t G, int64_t S, int64_t ldr, int32_t* H,
                     int64_t ldh) {
  MCPG_CHECK_HANDLE(h);
  mcpg::begin_call(h);
  if (G < 0 || S < 0) return mcpg::set_error(h, MCPG_ERR_SHAPE, "negative dimension");
  if (G == 0) return MCPG_OK;
  if (!R || !H) return mcpg::set_error(h, MCPG_ERR_INVALID_ARG, "null matrix pointer");
  if (ldr < S || ldh < S + 1) return mcpg::set_error(h, MCPG_ERR_INVALID_ARG, "leading dimension too small");
  if (S > 16 * 1024 * 1024) return mcpg::set_error(h, MCPG_ERR_UNSUPPORTED, "S too large");
  if (S + 1 > 16384) return mcpg::set_error(h, MCPG_ERR_UNSUPPORTED, "cumulative histogram supports S < 16384");
  if (ptrs != MCPG_HOST && ptrs != MCPG_DEVICE) return mcpg::set_error(h, MCPG_ERR_INVALID_ARG, "bad ptrs kind");
  MCPG_HIP(h, hipSetDevice(h->device));
  int st;
  const int32_t* dR = R;
  int32_t* dH = H;
  int64_t lr = ldr, lh = ldh;
  if (ptrs == MCPG_HOST) {
    mcpg::Sizer sz;
    sz.add<int32_t>(G * S);
    sz.add<int32_t>(G * (S + 1));
    void* base = mcpg::workspace(h, sz.off, &st);
    if (!base) return st;
    mcpg::Carver cv(base);
    int32_t* a = cv.take<int32_t>(G * S);
    dH = cv.take<int32_t>(G * (S + 1));
    if (S > 0 && (st = mcpg::copy_2d(h, a, S * 4, R, ldr * 4, S * 4, G, hipMemcpyHostToDevice))) return st;
    dR = a; lr = S; lh = S + 1;
  }
  if ((st = mcpg::reset_flags(h))) return st;
  {
    mcpg::LaunchScope ls(h, "cumhist_kernel");
    hipLaunchKernelGGL(cumhist_kernel, dim3((unsigned)G), dim3(256), (S + 1) * sizeof(int), h->stream, dR, lr, S, dH, lh,
                       h->d_flags, (int32_t*)nullptr);
  }
  if (ptrs == MCPG_HOST &&
      (st = mcpg::copy_2d(h, H, ldh * 4, dH, (S + 1) * 4, (S + 1) * 4, G, hipMemcpyDeviceToHost)))
    return st;
  if ((st = mcpg::finish(h, "mcpg_cumhist_i32"))) return st;
  return check_flags(h, "mcpg_cumhist_i32");
}

// The ranks' histograms, buckets and cell codes, then the pair kernel over the whole matrix
// (blk == nullptr), a block of it (blk = {row0, nrows, col0, ncols}) or the listed pairs (cells).
// MI64 non-null: the MI is written in double there instead of MI.
static int ap_run(mcpg_ctx* h, int ptrs, const int32_t* R, int64_t G, int64_t S, int64_t ldr, float* MI, int64_t ldmi,
                  const int32_t* px, const int32_t* py, int64_t npairs, int32_t max_leaves, int32_t* cells,
                  int32_t* nleaves, double* miv, const int64_t* blk = nullptr, double* MI64 = nullptr) {
  MCPG_HIP(h, hipSetDevice(h->device));
  int st;
  const bool cells_mode = npairs > 0;
  mcpg::Sizer sz;
  if (ptrs == MCPG_HOST) sz.add<int32_t>(G * S);
  sz.add<int32_t>(G * (S + 1));
  sz.add<int32_t>(G * S);
  sz.add<int32_t>(G);
  sz.add<uint16_t>(G * ap_ldz(S));
  const int64_t r0 = blk ? blk[0] : 0, nr = blk ? blk[1] : G, c0 = blk ? blk[2] : 0, nc = blk ? blk[3] : G;
  if (!cells_mode && ptrs == MCPG_HOST) {
    if (MI64) sz.add<double>(nr * nc);
    else sz.add<float>(nr * nc);
  }
  if (cells_mode) {
    sz.add<int32_t>(npairs); sz.add<int32_t>(npairs);
    sz.add<int32_t>(npairs * 3 * (int64_t)max_leaves); sz.add<int32_t>(npairs); sz.add<double>(npairs);
  }
  void* base = mcpg::workspace(h, sz.off, &st);
  if (!base) return st;
  mcpg::Carver cv(base);
  const int32_t* dR = R;
  int64_t lr = ldr;
  if (ptrs == MCPG_HOST) {
    int32_t* t = cv.take<int32_t>(G * S);
    if ((st = mcpg::copy_2d(h, t, S * 4, R, ldr * 4, S * 4, G, hipMemcpyHostToDevice))) return st;
    dR = t; lr = S;
  }
  int32_t* dH = cv.take<int32_t>(G * (S + 1));
  int32_t* dP = cv.take<int32_t>(G * S);
  int32_t* dId = cv.take<int32_t>(G);
  uint16_t* dZ = cv.take<uint16_t>(G * ap_ldz(S));
  float* dMI = MI;
  double* dMI64 = MI64;
  int64_t lm = ldmi;
  if (!cells_mode && ptrs == MCPG_HOST) {
    if (MI64) dMI64 = cv.take<double>(nr * nc);
    else dMI = cv.take<float>(nr * nc);
    lm = nc;
  }
  if ((st = mcpg::reset_flags(h))) return st;
  {
    mcpg::LaunchScope ls(h, "cumhist_kernel");
    hipLaunchKernelGGL(cumhist_kernel, dim3((unsigned)G), dim3(256), (S + 1) * sizeof(int), h->stream, dR, lr, S, dH,
                       S + 1, h->d_flags, dId);
  }
  // ranks index the histograms: refuse out-of-range ranks before the bucket and pair kernels read them
  if ((st = check_flags(h, "mcpg_ap_mi_i32"))) return st;
  {
    mcpg::LaunchScope ls(h, "ybucket_kernel");
    hipLaunchKernelGGL(ybucket_kernel, dim3((unsigned)G), dim3(256), S * sizeof(int), h->stream, dR, lr, S, dH, S + 1,
                       dP);
  }
  {
    mcpg::LaunchScope ls(h, "apcode_kernel");
    hipLaunchKernelGGL(apcode_kernel, dim3((unsigned)std::min<int64_t>(ceil_div(G * ap_ldz(S), 256), 65536)), dim3(256),
                       0, h->stream, dR, lr, G, S, dZ, ap_ldz(S));
  }
  APArgs a{};
  a.R = dR; a.ldr = lr; a.H = dH; a.ldh = S + 1; a.P = dP; a.G = G; a.S = S; a.MI = dMI; a.MI64 = dMI64; a.ldmi = lm;
  a.r0 = r0; a.nr = nr; a.c0 = c0; a.nc = nc; a.block = blk != nullptr;
  a.flags = h->d_flags;
  a.queue = reinterpret_cast<unsigned long long*>(h->d_flags + 8);   // zeroed by reset_flags
  a.ident = dId;
  a.Z = dZ;
  a.ldz = ap_ldz(S);
  int32_t *d_px = nullptr, *d_py = nullptr, *d_cells = nullptr, *d_nl = nullptr;
  double* d_mi = nullptr;
  if (cells_mode) {
    d_px = cv.take<int32_t>(npairs); d_py = cv.take<int32_t>(npairs);
    d_cells = cv.take<int32_t>(npairs * 3 * (int64_t)max_leaves); d_nl = cv.take<int32_t>(npairs);
    d_mi = cv.take<double>(npairs);
    MCPG_HIP(h, hipMemcpyAsync(d_px, px, npairs * 4, hipMemcpyHostToDevice, h->stream));
    MCPG_HIP(h, hipMemcpyAsync(d_py, py, npairs * 4, hipMemcpyHostToDevice, h->stream));
    a.px = d_px; a.py = d_py; a.npairs = npairs; a.max_leaves = max_leaves; a.cells = d_cells; a.nleaves = d_nl;
    a.mi_out = d_mi;
  }
  const int64_t work = cells_mode ? npairs : blk ? nr * nc : G * (G + 1) / 2;
  const int64_t per_wave = ap_wave_bytes(S);
  const int waves = (int)std::max<int64_t>(1, std::min<int64_t>(kMaxWaves, kLdsBudget / per_wave));
  const size_t lds = (size_t)waves * per_wave;
  int64_t blocks = std::max<int64_t>(1, std::min<int64_t>(ceil_div(work, waves), 256 * 64 / waves));
  if (!cells_mode) {   // queue mode: only as many workgroups as stay resident together
    int cus = 0, occ = 0;
    MCPG_HIP(h, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, h->device));
    MCPG_HIP(h, hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, reinterpret_cast<const void*>(&ap_mi_kernel<false>),
                                                             64 * waves, lds));
    blocks = std::max<int64_t>(1, std::min<int64_t>(ceil_div(work, (int64_t)waves * kChunk),
                                                    (int64_t)std::max(cus, 1) * std::max(occ, 1)));
  }
  {
    mcpg::LaunchScope ls(h, "ap_mi_kernel");
    if (cells_mode)
      hipLaunchKernelGGL(ap_mi_kernel<true>, dim3((unsigned)blocks), dim3(64 * waves), lds, h->stream, a,
                         (const int32_t*)dH);
    else
      hipLaunchKernelGGL(ap_mi_kernel<false>, dim3((unsigned)blocks), dim3(64 * waves), lds, h->stream, a,
                         (const int32_t*)dH);
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return mcpg::hip_error(h, e, "ap_mi_kernel launch");
  if (cells_mode) {
    MCPG_HIP(h, hipMemcpyAsync(cells, d_cells, npairs * 3 * (int64_t)max_leaves * 4, hipMemcpyDeviceToHost, h->stream));
    MCPG_HIP(h, hipMemcpyAsync(nleaves, d_nl, npairs * 4, hipMemcpyDeviceToHost, h->stream));
    MCPG_HIP(h, hipMemcpyAsync(miv, d_mi, npairs * 8, hipMemcpyDeviceToHost, h->stream));
  } else if (ptrs == MCPG_HOST && MI64) {
    if ((st = mcpg::copy_2d(h, MI64, ldmi * 8, dMI64, nc * 8, nc * 8, nr, hipMemcpyDeviceToHost))) return st;
  } else if (ptrs == MCPG_HOST) {
    if ((st = mcpg::copy_2d(h, MI, ldmi * 4, dMI, nc * 4, nc * 4, nr, hipMemcpyDeviceToHost))) return st;
  }
  if ((st = mcpg::finish(h, "mcpg_ap_mi_i32"))) return st;
  return check_flags(h, "mcpg_ap_mi_i32");
}

int mcpg_ap_mi_i32(mcpg_handle h, int ptrs, const int32_t* R, int64_t G, int64_t S, int64_t ldr, float* MI,
                   int64_t ldmi) {
  MCPG_CHECK_HANDLE(h);
  mcpg::begin_call(h);
  if (G < 0 || S < 0) return mcpg::set_error(h, MCPG_ERR_SHAPE, "negative dimension");
  if (G == 0) return MCPG_OK;
  if (S < 1) return mcpg::set_error(h, MCPG_ERR_SHAPE, "need at least one sample");
  if (S > kMaxS) return mcpg::set_error(h, MCPG_ERR_UNSUPPORTED, "adaptive-partitioning MI supports S <= 8191");
  if (!R || !MI) return mcpg::set_error(h, MCPG_ERR_INVALID_ARG, "null matrix pointer");
  if (ldr < S || ldmi < G) return mcpg::set_error(h, MCPG_ERR_INVALID_ARG, "leading dimension too small");
  if (ptrs != MCPG_HOST && ptrs != MCPG_DEVICE) return mcpg::set_error(h, MCPG_ERR_INVALID_ARG, "bad ptrs kind");
  return ap_run(h, ptrs, R, G, S, ldr, MI, ldmi, nullptr, nullptr, 0, 0, nullptr, nullptr, nullptr);
}

int mcpg_ap_mi_f64(mcpg_handle h, int ptrs, const int32_t* R, int64_t G, int64_t S, int64_t ldr, double* MI,
                   int64_t ldmi) {
  MCPG_CHECK_HANDLE(h);
  mcpg::begin_call(h);
  if (G < 0 || S < 0) return mcpg::set_error(h, MCPG_ERR_SHAPE, "negative dimension");
  if (G == 0) return MCPG_OK;
  if (S < 1) return mcpg::set_error(h, MCPG_ERR_SHAPE, "need at least one sample");
  if (S > kMaxS) return mcpg::set_error(h, MCPG_ERR_UNSUPPORTED, "adaptive-partitioning MI supports S <= 8191");
  if (!R || !MI) return mcpg::set_error(h, MCPG_ERR_INVALID_ARG, "null matrix pointer");
  if (ldr < S || ldmi < G) return mcpg::set_error(h, MCPG_ERR_INVALID_ARG, "leading dimension too small");
  if (ptrs != MCPG_HOST && ptrs != MCPG_DEVICE) return mcpg::set_error(h, MCPG_ERR_INVALID_ARG, "bad ptrs kind");
  return ap_run(h, ptrs, R, G, S, ldr, nullptr, ldmi, nullptr, nullptr, 0, 0, nullptr, nullptr, nullptr, nullptr, MI);
}

// the argument checks and run of both block entries (MI64 non-null: the double output)
static int ap_block_entry(mcpg_ctx* h, int ptrs, const int32_t* R, int64_t G, int64_t S, int64_t ldr, int64_t row0,
                          int64_t nrows, int64_t col0, int64_t ncols, float* MI, double* MI64, int64_t ldmi) {
  if (G < 0 || S < 0) return mcpg::set_error(h, MCPG_ERR_SHAPE, "negative dimension");
  if (row0 < 0 || nrows < 0 || col0 < 0 || ncols < 0 || row0 > G - nrows || col0 > G - ncols)
    return mcpg::set_error(h, MCPG_ERR_SHAPE, "block outside the G x G matrix");
  if (G == 0 || nrows == 0 || ncols == 0) return MCPG_OK;
  if (S < 1) return mcpg::set_error(h, MCPG_ERR_SHAPE, "need at least one sample");
  if (S > kMaxS) return mcpg::set_error(h, MCPG_ERR_UNSUPPORTED, "adaptive-partitioning MI supports S <= 8191");
  if (!R || (!MI && !MI64)) return mcpg::set_error(h, MCPG_ERR_INVALID_ARG, "null matrix pointer");
  if (ldr < S || ldmi < ncols) return mcpg::set_error(h, MCPG_ERR_INVALID_ARG, "leading dimension too small");
  if (ptrs != MCPG_HOST && ptrs != MCPG_DEVICE) return mcpg::set_error(h, MCPG_ERR_INVALID_ARG, "bad ptrs kind");
  const int64_t blk[4] = {row0, nrows, col0, ncols};
  return ap_run(h, ptrs, R, G, S, ldr, MI, ldmi, nullptr, nullptr, 0, 0, nullptr, nullptr, nullptr, blk, MI64);
}

int mcpg_ap_mi_block_f64(mcpg_handle h, int ptrs, const int32_t* R, int64_t G, int64_t S, int64_t ldr, int64_t row0,
                         int64_t nrows, int64_t col0, int64_t ncols, double* MI, int64_t ldmi) {
  MCPG_CHECK_HANDLE(h);
  mcpg::begin_call(h);
  return ap_block_entry(h, ptrs, R, G, S, ldr, row0, nrows, col0, ncols, nullptr, MI, ldmi);
}

int mcpg_ap_mi_block_i32(mcpg_handle h, int ptrs, const int32_t* R, int64_t G, int64_t S, int64_t ldr, int64_t row0,
                         int64_t nrows, int64_t col0, int64_t ncols, float* MI, int64_t ldmi) {
  MCPG_CHECK_HANDLE(h);
  mcpg::begin_call(h);
  return ap_block_entry(h, ptrs, R, G, S, ldr, row0, nrows, col0, ncols, MI, nullptr, ldmi);
}

int mcpg_ap_mi_cells_i32(mcpg_handle h, int ptrs, const int32_t* R, int64_t G, int64_t S, int64_t ldr,
                         const int32_t* px, const int32_t* py, int64_t npairs, int32_t max_leaves, int32_t* cells,
                         int32_t* nleaves, double* mi) {
  MCPG_CHECK_HANDLE(h);
  mcpg::begin_call(h);
  if (G < 0 || S < 0 || npairs < 0 || max_leaves < 0) return mcpg::set_error(h, MCPG_ERR_SHAPE, "negative dimension");
  if (G == 0 || npairs == 0) return MCPG_OK;
  if (S < 1) return mcpg::set_error(h, MCPG_ERR_SHAPE, "need at least one sample");
  if (S > kMaxS) return mcpg::set_error(h, MCPG_ERR_UNSUPPORTED, "adaptive-partitioning MI supports S <= 8191");
  if (!R || !px || !py || !nleaves || !mi || (max_leaves > 0 && !cells))
    return mcpg::set_error(h, MCPG_ERR_INVALID_ARG, "null pointer");
  if (ldr < S) return mcpg::set_error(h, MCPG_ERR_INVALID_ARG, "leading dimension too small");
  for (int64_t n = 0; n < npairs; ++n)
    if (px[n] < 0 || px[n] >= G || py[n] < 0 || py[n] >= G)
      return mcpg::set_error(h, MCPG_ERR_INVALID_ARG, "pair index out of range");
  if (ptrs != MCPG_HOST && ptrs != MCPG_DEVICE) return mcpg::set_error(h, MCPG_ERR_INVALID_ARG, "bad ptrs kind");
  return ap_run(h, ptrs, R, G, S, ldr, nullptr, 0, px, py, npairs, max_leaves, cells, nleaves, mi);
}

}  // extern "C"
