/*
 * mcpgpu.h -- C ABI of the MI355X (gfx950) implementation of MCPNet's two hot paths:
 * all-pairs mutual information (B-spline and adaptive partitioning) and the max-min
 * ("maximum capacity path") matrix products behind the MCP2/3/4 scores.
 *
 * Every entry point replaces one splash pattern call of the reference, specialised on the
 * reference's kernel functor; the comment above each declaration names the reference code
 * (path:line, relative to the MCPNet source tree) whose result it reproduces.
 *
 * Conventions
 *   - Matrices are row-major, rows = genes; leading dimensions (ld*) are in ELEMENTS.
 *   - `ptrs` says where the matrix pointers live: MCPG_HOST (pageable or pinned host memory;
 *     the call copies in, computes and copies out before returning; the MCP, maxmins and
 *     mcpnet entries copy each output out on a second stream while the next product, or row
 *     panel of the last product, runs, and return once every copy has ended) or MCPG_DEVICE
 *     (device memory on the handle's device; work is ordered on the handle's stream and the
 *     call returns after the stream has drained, so errors are reported synchronously).
 *   - Return value: MCPG_OK (0) or a negative mcpg_status; nothing is thrown across the ABI.
 *     mcpg_last_error() returns a message for the last failing call on a handle.
 *   - A handle owns a device, a HIP stream and a grow-only device workspace.  A handle is
 *     not thread-safe: use one per host thread (the reference copies its kernel functors
 *     per thread for the same reason).
 */
#ifndef MCPGPU_H_
#define MCPGPU_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mcpg_ctx* mcpg_handle;

enum mcpg_status {
  MCPG_OK = 0,
  MCPG_ERR_INVALID_ARG = -1,   /* null pointer, bad enum, bins < order, ld too small ... */
  MCPG_ERR_SHAPE = -2,         /* dimension mismatch (the reference asserts, BSplineMI.hpp:332, :493) */
  MCPG_ERR_NONFINITE = -3,     /* NaN/Inf in an input that must be finite */
  MCPG_ERR_OUT_OF_MEMORY = -4, /* device allocation failed */
  MCPG_ERR_HIP = -5,           /* any other HIP runtime error (text in mcpg_last_error) */
  MCPG_ERR_NO_DEVICE = -6,     /* no HIP device / bad device ordinal */
  MCPG_ERR_UNSUPPORTED = -7    /* size outside what the kernels support (e.g. S too large) */
};

enum mcpg_ptr_kind { MCPG_HOST = 0, MCPG_DEVICE = 1 };

/* Arithmetic of the B-spline joint-table contraction. */
enum mcpg_precision {
  MCPG_FP32 = 0,   /* FP32 weights, FP32 MFMA (v_mfma_f32_16x16x4_f32) */
  MCPG_BF16 = 1,   /* BF16 weights, BF16 MFMA, FP32 accumulation */
  MCPG_BF16X3 = 2, /* hi+lo BF16 split, 3 MFMA products (hi*hi + hi*lo + lo*hi), FP32 accumulation */
  MCPG_I8X3 = 3,   /* two int8 digits per weight, w ~= (h + l/254)/127; products h*h and h*l + l*h
                      on the int8 MFMA (v_mfma_i32_16x16x64_i8) with exact int32 sums;
                      S <= 66560.  Not in the reference (its MI is FP32/FP64 on the CPU). */
  MCPG_FP64 = 4    /* the reference's double DataType: scaling, basis, marginals and joint tables
                      in double (v_mfma_f64_16x16x4_f64).  Taken by the *_f64 MI entry points
                      below and by mcpg_mcpnet_*; the *_f32 MI entry points reject it. */
};

/* ------------------------------------------------------------------ handle */
int mcpg_create(mcpg_handle* h, int device);
int mcpg_destroy(mcpg_handle h);
const char* mcpg_status_string(int status);
const char* mcpg_last_error(mcpg_handle h);
int mcpg_synchronize(mcpg_handle h);
/* The hipStream_t the handle launches on (returned as void*), for callers that time or
 * order work around the library's kernels. */
int mcpg_get_stream(mcpg_handle h, void** stream);
/* Number of kernel launches the most recent call made and, for the dominant kernel of that
 * call, its name (a static string) -- used by profiling harnesses. */
int mcpg_last_launch_info(mcpg_handle h, int* launches, const char** dominant_kernel);
/* Profiling (no counterpart in the reference, whose stage timers are host wall clocks,
 * mcpnet.cpp:289-296): when enabled, every kernel launch of a call is bracketed by HIP events on
 * the handle's stream; mcpg_last_timings returns the kernel names (static strings) and durations
 * (ms) of the most recent call's launches, up to max_entries, and their total count. */
int mcpg_set_profiling(mcpg_handle h, int enable);
int mcpg_last_timings(mcpg_handle h, int max_entries, const char** names, float* ms, int* count);

/* ------------------------------------------------------------------ B-spline MI */

/* MinMaxScale<float> row transform: y = (x - min) * (1/(max-min)), a constant row is copied.
 * Replaces GlobalTransform<..., mcp::kernel::MinMaxScale<T>> (include/mcp/correlation/BSplineMI.hpp:75-116;
 * called at apps/pipeline/mcpnet_helper.hpp:60-64, apps/correlation/mi.cpp:198-204). */
int mcpg_minmax_scale_f32(mcpg_handle h, int ptrs, const float* X, int64_t G, int64_t S, int64_t ldx,
                          float* Y, int64_t ldy);

/* BSplineWeightsKernel<float> row transform on already-scaled rows: W[g][b*S+s] = B_{b,order}(x_gs)
 * (bin-major, the reference layout) and W[g][bins*S] = marginal entropy H_g (computed in FP64,
 * stored as float).  ldw >= bins*S+1.
 * Replaces Transform<..., mcp::kernel::BSplineWeightsKernel<T>> (BSplineMI.hpp:119-274;
 * mcpnet_helper.hpp:70-80, mi.cpp:214-222). */
int mcpg_bspline_weights_f32(mcpg_handle h, int ptrs, const float* Xs, int64_t G, int64_t S, int64_t ldx,
                             int bins, int order, float* W, int64_t ldw);

/* All-pairs B-spline MI from reference-layout weight rows (the output of
 * mcpg_bspline_weights_f32): MI[i][j] = W[i][bins*S] + W[j][bins*S] - H(i,j) for all i, j,
 * with H(i,j) = -sum_{b,c} p log2 p, p = (1/S) <W_i,b , W_j,c>.  The output is exactly symmetric.
 * Replaces InnerProduct<..., mcp::kernel::BSplineMIKernel<T>, ..., true> (BSplineMI.hpp:276-336,
 * :492-499; mcpnet_helper.hpp:90-98; the commented-out call at mi.cpp:251). */
int mcpg_bspline_mi_weights_f32(mcpg_handle h, int ptrs, const float* W, int64_t G, int64_t S, int64_t ldw,
                                int bins, int precision, float* MI, int64_t ldmi);

/* Fused compute_mi (apps/pipeline/mcpnet_helper.hpp:56-99): min-max scale, B-spline weights with
 * FP64 marginal entropies, all-pairs MI with the chosen contraction precision; the G x G MI
 * matrix is exactly symmetric.  Weight panels never leave the device. */
int mcpg_bspline_mi_f32(mcpg_handle h, int ptrs, const float* X, int64_t G, int64_t S, int64_t ldx,
                        int bins, int order, int precision, float* MI, int64_t ldmi);

/* The same for one block of the G x G MI matrix: MI (nrows x ncols, ldmi >= ncols) receives rows
 * row0 .. row0+nrows-1 and columns col0 .. col0+ncols-1, bit-identical to those entries of
 * mcpg_bspline_mi_f32.  This is the unit a rank of the reference's distributed run fills: the
 * InnerProduct of mcpnet_helper.hpp:90-98 writes the rank's row partition of `dmi` (all G
 * columns) from the allgathered weights (:84).  Weights and marginals are computed for all G
 * genes; only the tiles that cover the block are launched. */
int mcpg_bspline_mi_block_f32(mcpg_handle h, int ptrs, const float* X, int64_t G, int64_t S, int64_t ldx,
                              int bins, int order, int precision, int64_t row0, int64_t nrows, int64_t col0,
                              int64_t ncols, float* MI, int64_t ldmi);

/* FP64 forms of the four B-spline entry points above, for the reference's double DataType
 * (mi.cpp:300-304 and mcpnet.cpp run<double> unless --single): MinMaxScale<double>,
 * BSplineWeightsKernel<double> and BSplineMIKernel<double> (BSplineMI.hpp:75-116, :119-274,
 * :339-500).  Scale 1.0/(max-min) and x' = scale*(x-min) in double, knots of
 * buildKnotVector<double>, the recursive basis in double, p = (1/S)<W_i,b , W_j,c> with the dot
 * products on the FP64 MFMA, MI = H_i + H_j - H_ij.  The weight rows of mcpg_bspline_weights_f64
 * hold H in double at [bins*S].  MI results agree with the reference's double arithmetic up to
 * the summation order of the dot products. */
int mcpg_minmax_scale_f64(mcpg_handle h, int ptrs, const double* X, int64_t G, int64_t S, int64_t ldx,
                          double* Y, int64_t ldy);
int mcpg_bspline_weights_f64(mcpg_handle h, int ptrs, const double* Xs, int64_t G, int64_t S, int64_t ldx,
                             int bins, int order, double* W, int64_t ldw);
int mcpg_bspline_mi_weights_f64(mcpg_handle h, int ptrs, const double* W, int64_t G, int64_t S, int64_t ldw,
                                int bins, double* MI, int64_t ldmi);
int mcpg_bspline_mi_f64(mcpg_handle h, int ptrs, const double* X, int64_t G, int64_t S, int64_t ldx,
                        int bins, int order, double* MI, int64_t ldmi);
/* The block of mcpg_bspline_mi_f64, as mcpg_bspline_mi_block_f32 is of mcpg_bspline_mi_f32 (the
 * double DataType's row partition of mcpnet_helper.hpp:90-98), bit-identical to those entries. */
int mcpg_bspline_mi_block_f64(mcpg_handle h, int ptrs, const double* X, int64_t G, int64_t S, int64_t ldx,
                              int bins, int order, int64_t row0, int64_t nrows, int64_t col0, int64_t ncols,
                              double* MI, int64_t ldmi);

/* Adjoint of the B-spline MI (no counterpart in the reference, which has no gradients):
 * for L = sum_n m[n] * MI(pi[n], pj[n]) (pairs with pi != pj; the MI is the MCPG_FP32 forward),
 * writes dX[g][s] = dL/dX[g][s] through the min-max scaling (sub-gradient at the first arg-min /
 * arg-max), the B-spline basis derivative and the FP32 joint tables, which are recomputed from
 * X rather than stored.  pair arrays are in host memory; X and dX follow `ptrs`.
 * If L_out is non-null it receives L (FP64 accumulation of the FP32 forward). */
int mcpg_bspline_mi_backward_f32(mcpg_handle h, int ptrs, const float* X, int64_t G, int64_t S, int64_t ldx,
                                 int bins, int order, const int32_t* pi, const int32_t* pj, const float* m,
                                 int64_t npairs, float* dX, int64_t lddx, double* L_out);

/* ------------------------------------------------------------------ adaptive-partitioning MI */

/* add_noise: out = in + U[0, 1e-5) jitter from a counter-based SplitMix64 stream
 * (element (g,s) draws counter g*S+s of stream `seed`).  Replaces add_noise
 * (include/mcp/correlation/AdaptivePartitioningMI.hpp:40-114; mi.cpp:125-132), whose random
 * stream depends on the thread count and is not reproduced. */
int mcpg_add_noise_f32(mcpg_handle h, int ptrs, const float* X, int64_t G, int64_t S, int64_t ldx,
                       uint64_t seed, float* Y, int64_t ldy);

/* Descending rank per row (largest value -> rank 0), ties broken by ascending column index.
 * Replaces Transform<..., splash::kernel::Rank<T,size_t,0,false>> (mi.cpp:141-147,
 * mcpnet_helper.hpp:127-133). */
int mcpg_rank_f32(mcpg_handle h, int ptrs, const float* X, int64_t G, int64_t S, int64_t ldx,
                  int32_t* R, int64_t ldr);

/* IntegralCumulativeHistogramKernel<int,0>: H[g][i] = #{s : R[g][s] < i}, i = 0..S (ldh >= S+1).
 * (AdaptivePartitioningMI.hpp:414-442; mi.cpp:156-162). */
int mcpg_cumhist_i32(mcpg_handle h, int ptrs, const int32_t* R, int64_t G, int64_t S, int64_t ldr,
                     int32_t* H, int64_t ldh);

/* All-pairs adaptive-partitioning MI on rank rows (values in [0,S)): replaces
 * InnerProduct<..., AdaptivePartitionRankMIKernel2<IT,T,0>, ..., true> (AdaptivePartitioningMI.hpp:445-647;
 * mi.cpp:170-177, mcpnet_helper.hpp:156-162).  MI[i][j] for i <= j is computed as kernel(x=i, y=j)
 * and mirrored.  Integer cell counts follow the reference exactly; the per-pair sum is FP64 in
 * the reference's LIFO window order. */
int mcpg_ap_mi_i32(mcpg_handle h, int ptrs, const int32_t* R, int64_t G, int64_t S, int64_t ldr,
                   float* MI, int64_t ldmi);

/* The same with the MI kept in double: AdaptivePartitionRankMIKernel2<IT, double, 0>, the
 * reference's double DataType, whose result mi/S + log2 S (AdaptivePartitioningMI.hpp:644) is not
 * rounded to float.  Bit-identical to mcpg_ap_mi_i32 before that rounding. */
int mcpg_ap_mi_f64(mcpg_handle h, int ptrs, const int32_t* R, int64_t G, int64_t S, int64_t ldr,
                   double* MI, int64_t ldmi);

/* One block of the same matrix (rows row0 .. row0+nrows-1, columns col0 .. col0+ncols-1 into MI,
 * ldmi >= ncols), bit-identical to those entries of mcpg_ap_mi_i32: each entry (i, j) is
 * kernel(x=min(i,j), y=max(i,j)).  The row-partition unit of mi.cpp:170-177 and
 * mcpnet_helper.hpp:156-162, where each rank's InnerProduct fills its rows of the distributed MI. */
int mcpg_ap_mi_block_i32(mcpg_handle h, int ptrs, const int32_t* R, int64_t G, int64_t S, int64_t ldr,
                         int64_t row0, int64_t nrows, int64_t col0, int64_t ncols, float* MI, int64_t ldmi);
/* The block of mcpg_ap_mi_f64 (the double DataType's row partition). */
int mcpg_ap_mi_block_f64(mcpg_handle h, int ptrs, const int32_t* R, int64_t G, int64_t S, int64_t ldr,
                         int64_t row0, int64_t nrows, int64_t col0, int64_t ncols, double* MI, int64_t ldmi);

/* Debug/parity view of the same kernel: for each requested pair (px[n], py[n]) the leaf windows
 * in the order their terms enter the FP64 sum, as (nxy, nx, ny) int32 triples, at most
 * max_leaves per pair (cells[n*max_leaves*3 + 3*l + {0,1,2}]); nleaves[n] gets the true count
 * (which may exceed max_leaves) and mi[n] the pair's MI.  Pair arrays and outputs are host memory. */
int mcpg_ap_mi_cells_i32(mcpg_handle h, int ptrs, const int32_t* R, int64_t G, int64_t S, int64_t ldr,
                         const int32_t* px, const int32_t* py, int64_t npairs, int32_t max_leaves,
                         int32_t* cells, int32_t* nleaves, double* mi);

/* ------------------------------------------------------------------ max-min (MCP) */

/* C[i][j] = max_k min(A[i][k], B[j][k])  (row . row, like splash InnerProduct), i < M, j < N, k < K.
 * kmask (nullable, device/host per `ptrs`, length K): MASKED kernel, the term becomes
 * min(kmask[k], A[i][k], B[j][k]).  Empty K gives -inf (f32) / -DBL_MAX (f64) as in the reference.
 * symmetric != 0 requires A == B (same pointer) and M == N and computes only j >= i, mirroring.
 * Replaces InnerProduct<..., mcp::kernel::mcp2_maxmin_kernel<T,MASKED,true>, ...>
 * (include/mcp/filter/mcp.hpp:72-93 -> include/mcp/filter/simd_maxmin.hpp:30-46, :299-450).
 * Exact: bit-identical to the reference for inputs without NaN or -0. */
int mcpg_maxmin_f32(mcpg_handle h, int ptrs, const float* A, int64_t lda, const float* B, int64_t ldb,
                    const float* kmask, int64_t M, int64_t N, int64_t K, int symmetric, float* C, int64_t ldc);
int mcpg_maxmin_f64(mcpg_handle h, int ptrs, const double* A, int64_t lda, const double* B, int64_t ldb,
                    const double* kmask, int64_t M, int64_t N, int64_t K, int symmetric, double* C, int64_t ldc);

/* ratio_kernel<T,T> (include/mcp/transform/combine.hpp:210-222; compute_ratio mcpnet_helper.hpp:172-184):
 * out = |a| < eps ? 0 : |m| < eps ? T_MAX : (T)((double)a / (double)m), eps = numeric_limits<T>::epsilon(). */
int mcpg_ratio_f32(mcpg_handle h, int ptrs, const float* A, int64_t lda, const float* Mm, int64_t ldm,
                   int64_t M, int64_t N, float* out, int64_t ldo);
int mcpg_ratio_f64(mcpg_handle h, int ptrs, const double* A, int64_t lda, const double* Mm, int64_t ldm,
                   int64_t M, int64_t N, double* out, int64_t ldo);

/* max_kernel<T> (include/mcp/transform/combine.hpp:42-58; used for max(M, M^T) at apps/filter/mcp.cpp:463-474):
 * out = max(A, B) elementwise. */
int mcpg_elementwise_max_f32(mcpg_handle h, int ptrs, const float* A, int64_t lda, const float* B, int64_t ldb,
                             int64_t M, int64_t N, float* out, int64_t ldo);
int mcpg_elementwise_max_f64(mcpg_handle h, int ptrs, const double* A, int64_t lda, const double* B, int64_t ldb,
                             int64_t M, int64_t N, double* out, int64_t ldo);

/* compute_mcp{2,3,4} (apps/pipeline/mcpnet_helper.hpp:198-205, :270-277, :357-364) on an N x N
 * matrix A (MI with its diagonal already set): k = 2: M = A(x)A; k = 3: M = A(x)(A(x)A);
 * k = 4: M = (A(x)A)(x)(A(x)A); out = ratio(A, M).  maxmin_out (nullable) receives M (the
 * reference's -x/--maxmin-output).  Intermediates stay on the device; the last product fuses
 * the ratio.  Products whose result is exactly symmetric are computed on one triangle. */
int mcpg_mcp_f32(mcpg_handle h, int ptrs, const float* A, int64_t lda, int64_t N, int k,
                 float* maxmin_out, int64_t ldm, float* mcp_out, int64_t ldo);
int mcpg_mcp_f64(mcpg_handle h, int ptrs, const double* A, int64_t lda, int64_t N, int k,
                 double* maxmin_out, int64_t ldm, double* mcp_out, int64_t ldo);

/* Allocation-free MCP2/3/4.  No counterpart in the reference: this entry point is for callers
 * that cannot allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  It computes what mcpg_mcp_f32 computes, with the same
 * kernels and the same bits, on DEVICE pointers only (there is no ptrs argument), and allocates
 * nothing: the order keys of A and of M1, the triangle tile order and the split-tail key blocks
 * live in a static 64 MiB device array of the library, which each GPU the library runs on holds
 * until the process ends.  Limit: the keys take 4 * ceil(N/128)*128 * ceil(N/32)*32 bytes per
 * matrix (one matrix for k = 2, two for k = 3, 4), so N <= 4064 for k = 2 and N <= 2848 for
 * k = 3, 4; a larger N returns MCPG_ERR_UNSUPPORTED.  Every handle on one GPU shares that GPU's
 * array and each call overwrites it: calls that overlap on one GPU (two host threads, two
 * handles) race and must be serialised by the caller. */
int mcpg_mcp_f32_noalloc(mcpg_handle h, const float* A, int64_t lda, int64_t N, int k,
                         float* maxmin_out, int64_t ldm, float* mcp_out, int64_t ldo);

/* The mcpnet stage chain (apps/pipeline/mcpnet.cpp:286-515 with compute_mi / compute_ap_mi, the
 * diagonal, filter_local_mat_rows_by_tfs and compute_mcp{2,3,4}[_tfs], mcpnet_helper.hpp:56-435)
 * in one call, every intermediate in device memory: with MCPG_HOST pointers the only transfers
 * are X in and the requested outputs out.
 *   mi_method: 0 = B-spline MI (bins, order, precision as mcpg_bspline_mi_f32, or MCPG_FP64 as
 *     mcpg_bspline_mi_f64 on double values of X), 1 = adaptive
 *     partitioning (noise_seed, as mcpg_add_noise_f32 -> mcpg_rank_f32 -> mcpg_ap_mi_i32),
 *     -1 = X already is the G x G MI matrix (S ignored, ldx >= G).
 *   diagonal >= 0 sets the MI diagonal (mcpnet.cpp:299-308); a negative value keeps it.
 *   tfs (nullable, length G, memory per `ptrs`): > 0 marks a TF.  The MCP outputs then hold the
 *     T TF rows (T = number of TFs, in gene order) and order k follows compute_mcp{k}_tfs with
 *     tf_gene_transition k - 1; the tfs values themselves are the masked kernel's kmask.
 *   mi (nullable): the whole G x G MI after the diagonal is set (the ensemble stage needs all of
 *     it; its TF rows are what mcpnet writes in TF mode).
 *   mcp2 / mcp3 / mcp4 (nullable): the MCP scores of that order, G or T rows by G columns.  Without
 *     TFs, two or more orders share their max-min matrices (A (x) A once, as compute_maxmins,
 *     mcpnet_helper.hpp:436-460); the scores equal the per-order mcpg_mcp_* results bitwise.
 * The f64 form with precision MCPG_FP64 computes the MI in double (the reference's double
 * DataType); with the other precisions in FP32-class arithmetic from FP32 values of X, widened
 * before the diagonal.  Its max-min products and ratios run in FP64, as the stage entry points'.
 * The f32 form with MCPG_FP64 rounds the double MI to float. */
int mcpg_mcpnet_f32(mcpg_handle h, int ptrs, const float* X, int64_t G, int64_t S, int64_t ldx, int mi_method, int bins,
                    int order, int precision, uint64_t noise_seed, double diagonal, const float* tfs, float* mi,
                    int64_t ldmi, float* mcp2, int64_t ld2, float* mcp3, int64_t ld3, float* mcp4, int64_t ld4);
int mcpg_mcpnet_f64(mcpg_handle h, int ptrs, const double* X, int64_t G, int64_t S, int64_t ldx, int mi_method,
                    int bins, int order, int precision, uint64_t noise_seed, double diagonal, const double* tfs,
                    double* mi, int64_t ldmi, double* mcp2, int64_t ld2, double* mcp3, int64_t ld3, double* mcp4,
                    int64_t ld4);

/* compute_maxmins (apps/pipeline/mcpnet_helper.hpp:436-460), the three products the ensemble
 * MCP combines (apps/pipeline/mcpnet.cpp:546-553): M1 = A(x)A, M2 = A(x)M1, M3 = M1(x)M1 of the
 * N x N matrix A.  Each output is written only when its pointer is non-null; M1 stays on the
 * device between the products.  Exact, like mcpg_maxmin. */
int mcpg_maxmins_f32(mcpg_handle h, int ptrs, const float* A, int64_t lda, int64_t N, float* M1, int64_t ld1,
                     float* M2, int64_t ld2, float* M3, int64_t ld3);
int mcpg_maxmins_f64(mcpg_handle h, int ptrs, const double* A, int64_t lda, int64_t N, double* M1, int64_t ld1,
                     double* M2, int64_t ld2, double* M3, int64_t ld3);

/* ------------------------------------------------------------------ ensemble and filters */

/* compute_combo (apps/pipeline/mcpnet_helper.hpp:509-571; called per coefficient row at
 * apps/pipeline/mcpnet.cpp:589-592): with (a, b, c, d) = coeffs and s = a*MI + b*M1 + c*M2 + d*M3
 * accumulated in double (a coefficient of 1 adds the value unscaled, 0 skips the term and its
 * matrix, which may then be NULL), out = |MI| < eps ? 0 : |s| < eps ? T_MAX : (T)(MI / s).
 * Bit-identical to an IEEE evaluation of that expression (no contraction). */
int mcpg_combo_f32(mcpg_handle h, int ptrs, const float* MI, int64_t ldmi, const float* M1, int64_t ld1,
                   const float* M2, int64_t ld2, const float* M3, int64_t ld3, int64_t M, int64_t N,
                   const float* coeffs, float* out, int64_t ldo);
int mcpg_combo_f64(mcpg_handle h, int ptrs, const double* MI, int64_t ldmi, const double* M1, int64_t ld1,
                   const double* M2, int64_t ld2, const double* M3, int64_t ld3, int64_t M, int64_t N,
                   const double* coeffs, double* out, int64_t ldo);

/* compute_stouffer (apps/pipeline/mcpnet_helper.hpp:574-584): per-row mean and unbiased standard
 * deviation (splash::kernel::GaussianParamsExclude1<T,T,true>; exclude_ones != 0 leaves entries
 * equal to 1 out of both moments), then stouffer_vector_kernel (include/mcp/transform/stouffer.hpp:56-80):
 * out[i][j] = (z_i + z_j) / sqrt(2), z_r = (X[i][j] - mean_r) / sd_r (0 when sd_r < eps).
 * The moments are FP64; the splash reduction they replace is not in the reference tree. */
int mcpg_stouffer_f32(mcpg_handle h, int ptrs, const float* X, int64_t ldx, int64_t N, int exclude_ones,
                      float* out, int64_t ldo);
int mcpg_stouffer_f64(mcpg_handle h, int ptrs, const double* X, int64_t ldx, int64_t N, int exclude_ones,
                      double* out, int64_t ldo);

/* compute_stouffer_tf (apps/pipeline/mcpnet_helper.hpp:587-638; apps/transform/transform.cpp:160-200)
 * on an M x N TFs-to-genes score matrix: Transform<M, StandardScore<T, T, false>, M> z-scores of
 * every row (a TF over its genes) and of every column (a gene over the TFs, through the
 * transpose), combined by BinaryOp<M, M, zscored_stouffer_kernel<T>, M>
 * (include/mcp/transform/stouffer.hpp:84-99): out = (zr + zc) * sqrt(1/2).  z = (x - mean) / sd
 * with the population standard deviation, moments in FP64; a row or column with sd < eps(T)
 * scores 0. */
int mcpg_stouffer_tf_f32(mcpg_handle h, int ptrs, const float* X, int64_t ldx, int64_t M, int64_t N, float* out,
                         int64_t ldo);
int mcpg_stouffer_tf_f64(mcpg_handle h, int ptrs, const double* X, int64_t ldx, int64_t M, int64_t N, double* out,
                         int64_t ldo);

/* DPI filter (apps/filter/dpi.cpp:236-300): the mask of InnerProduct<..., dpi_kernel<T>> or, with
 * a TF list, dpi_tf_kernel<T> (include/mcp/filter/dpi.hpp:42-110, EQUALITY_EXCLUSION), applied with
 * mask<T, true> (include/mcp/filter/threshold.hpp:34-54): out[x][z] = 0 when x == z or some gene
 * y != x, z (a TF when exactly one of x, z is a TF) has A[x][y] >= i_xz and A[z][y] >= i_xz,
 * i_xz = A[x][z] / (1 - tolerance); else A[x][z].  tfs (nullable, length N): > 0 marks a TF.
 * mask (nullable, one byte per entry) receives the decision itself (1 = remove), the boolean
 * matrix the reference's InnerProduct produces; out and mask may each be NULL, not both.
 * removed (nullable) receives the number of entries of out that changed.  The max over y is a
 * max-min product with the excluded entries at -inf, so the decision is exact. */
int mcpg_dpi_f32(mcpg_handle h, int ptrs, const float* A, int64_t lda, int64_t N, float tolerance, const float* tfs,
                 float* out, int64_t ldo, uint8_t* mask, int64_t ldmask, int64_t* removed);
int mcpg_dpi_f64(mcpg_handle h, int ptrs, const double* A, int64_t lda, int64_t N, double tolerance,
                 const double* tfs, double* out, int64_t ldo, uint8_t* mask, int64_t ldmask, int64_t* removed);

#ifdef __cplusplus
}
#endif

#endif  /* MCPGPU_H_ */
