// Ensemble MCP epilogues on the score matrices: the linear-combination ratio of compute_combo,
// the Stouffer combination of compute_stouffer (N x N) and of compute_stouffer_tf (TFs x genes).
//
// Both are HBM-bound elementwise / row-reduction passes.  Arithmetic follows the reference
// expression by expression (compiled without contraction, so no FMA rewrites the products).
#include <cfloat>
#include <cmath>

#include "internal.h"

#pragma clang fp contract(off)

namespace {

using mcpg::ceil_div;

template <typename T>
struct Lim;
template <>
struct Lim<float> {
  static constexpr float eps = FLT_EPSILON;
  static constexpr float big = FLT_MAX;
};
template <>
struct Lim<double> {
  static constexpr double eps = DBL_EPSILON;
  static constexpr double big = DBL_MAX;
};

// compute_combo (apps/pipeline/mcpnet_helper.hpp:509-571): s = a*mi + b*m1 + c*m2 + d*m3 accumulated
// in double (a coefficient of 1 adds the value unscaled, 0 skips the term), then
// out = |mi| < eps ? 0 : (|s| < eps ? max : mi / s) with the division in double.
template <typename T>
__global__ void combo_kernel(const T* __restrict__ MI, int64_t ldmi, const T* __restrict__ M1, int64_t ld1,
                             const T* __restrict__ M2, int64_t ld2, const T* __restrict__ M3, int64_t ld3, int64_t M,
                             int64_t N, double a, double b, double c, double d, T* __restrict__ O, int64_t ldo) {
  const int64_t total = M * N;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = t / N, j = t % N;
    const T mi = MI[i * ldmi + j];
    double s = 0.0;
    if (a == 1.0) s += mi; else if (a != 0.0) s += a * mi;
    if (b == 1.0) s += M1[i * ld1 + j]; else if (b != 0.0) s += b * M1[i * ld1 + j];
    if (c == 1.0) s += M2[i * ld2 + j]; else if (c != 0.0) s += c * M2[i * ld2 + j];
    if (d == 1.0) s += M3[i * ld3 + j]; else if (d != 0.0) s += d * M3[i * ld3 + j];
    O[i * ldo + j] = (fabs((double)mi) < (double)Lim<T>::eps) ? (T)0
                     : (fabs(s) < (double)Lim<T>::eps) ? Lim<T>::big
                                                       : (T)((double)mi / s);
  }
}

// Per-row mean and unbiased standard deviation (n - 1 in the variance): the Gaussian row
// parameters compute_stouffer reduces each row to (splash::kernel::GaussianParamsExclude1<T, T,
// true>, mcpnet_helper.hpp:580; "degree of freedom 1", apps/transform/transform.cpp:216-218).
// With exclude_ones, entries equal to 1 (a ratio whose direct edge is its own widest path, or a
// self-correlation) are left out of both moments.  Two passes over the row in FP64, one block per
// row.
template <typename T>
__global__ __launch_bounds__(256) void row_stats_kernel(const T* __restrict__ X, int64_t ldx, int64_t N,
                                                        int exclude_ones, T* __restrict__ mean, T* __restrict__ sd) {
  const int64_t r = blockIdx.x;
  const T* x = X + r * ldx;
  __shared__ double sh[4];
  auto block_sum = [&](double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    return sh[0] + sh[1] + sh[2] + sh[3];
  };
  double s = 0.0, c = 0.0;
  for (int64_t j = threadIdx.x; j < N; j += 256) {
    if (exclude_ones && x[j] == (T)1) continue;
    s += (double)x[j];
    c += 1.0;
  }
  const double n = block_sum(c);
  const double mu = n > 0.0 ? block_sum(s) / n : 0.0;
  double q = 0.0;
  for (int64_t j = threadIdx.x; j < N; j += 256) {
    if (exclude_ones && x[j] == (T)1) continue;
    const double dv = (double)x[j] - mu;
    q += dv * dv;
  }
  const double var = n > 1.0 ? block_sum(q) / (n - 1.0) : 0.0;
  if (threadIdx.x == 0) {
    mean[r] = (T)mu;
    sd[r] = (T)sqrt(var);
  }
}

// stouffer_vector_kernel (include/mcp/transform/stouffer.hpp:56-80) for row i:
//   zi = (x_ij - mean_i) * denom_i, denom_i = |sd_i| < eps ? 0 : 1 / sd_i  (denominator in double,
//   stored in T); zj = |sd_j| < eps ? 0 : (x_ij - mean_j) / sd_j; out = (zi + zj) * (T)sqrt(1/2).
template <typename T>
__global__ void stouffer_kernel(const T* __restrict__ X, int64_t ldx, int64_t M, int64_t N, const T* __restrict__ mean,
                                const T* __restrict__ sd, T* __restrict__ O, int64_t ldo) {
  const T isqrt2 = (T)0.70710678118654752440084436210485L;
  const int64_t total = M * N;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = t / N, j = t % N;
    const T x = X[i * ldx + j];
    const T denom = (fabs(sd[i]) < Lim<T>::eps) ? (T)0 : (T)(1.0 / (double)sd[i]);
    const T zi = (T)((x - mean[i]) * denom);
    const T zj = (fabs(sd[j]) < Lim<T>::eps) ? (T)0 : (T)((x - mean[j]) / sd[j]);
    O[i * ldo + j] = (zi + zj) * isqrt2;
  }
}

// StandardScore<T, T, false> moments (splash; used by compute_stouffer_tf,
// apps/pipeline/mcpnet_helper.hpp:591-613): mean and population standard deviation (n in the
// variance; with them the reference's Pearson correlation, a dot product of z-scores divided by
// the count at include/mcp/correlation/Pearson.hpp:34-36, is exact) of each row, in FP64, one
// block per row, two passes.
template <typename T>
__global__ __launch_bounds__(256) void zrow_moments_kernel(const T* __restrict__ X, int64_t ldx, int64_t N,
                                                           double* __restrict__ mean, double* __restrict__ sd) {
  const int64_t r = blockIdx.x;
  const T* x = X + r * ldx;
  __shared__ double sh[4];
  auto block_sum = [&](double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    return sh[0] + sh[1] + sh[2] + sh[3];
  };
  double s = 0.0;
  for (int64_t j = threadIdx.x; j < N; j += 256) s += (double)x[j];
  const double mu = block_sum(s) / (double)N;
  double q = 0.0;
  for (int64_t j = threadIdx.x; j < N; j += 256) {
    const double dv = (double)x[j] - mu;
    q += dv * dv;
  }
  const double var = block_sum(q) / (double)N;
  if (threadIdx.x == 0) {
    mean[r] = mu;
    sd[r] = sqrt(var);
  }
}

// The same moments down each column of an M x N matrix: 64 columns per block (one per lane, so
// each wave reads 64 consecutive elements of a row), the 4 waves split the rows.
template <typename T>
__global__ __launch_bounds__(256) void zcol_moments_kernel(const T* __restrict__ X, int64_t ldx, int64_t M, int64_t N,
                                                           double* __restrict__ mean, double* __restrict__ sd) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int64_t c = (int64_t)blockIdx.x * 64 + lane;
  __shared__ double sh[4][64];
  double s = 0.0;
  if (c < N)
    for (int64_t r = w; r < M; r += 4) s += (double)X[r * ldx + c];
  sh[w][lane] = s;
  __syncthreads();
  const double mu = (sh[0][lane] + sh[1][lane] + sh[2][lane] + sh[3][lane]) / (double)M;
  __syncthreads();
  double q = 0.0;
  if (c < N)
    for (int64_t r = w; r < M; r += 4) {
      const double dv = (double)X[r * ldx + c] - mu;
      q += dv * dv;
    }
  sh[w][lane] = q;
  __syncthreads();
  if (w == 0 && c < N) {
    mean[c] = mu;
    sd[c] = sqrt((sh[0][lane] + sh[1][lane] + sh[2][lane] + sh[3][lane]) / (double)M);
  }
}

// compute_stouffer_tf's combination (zscored_stouffer_kernel, include/mcp/transform/stouffer.hpp:84-99)
// of the row z-score (TF normalised over its genes) and the column z-score (gene normalised over
// the TFs): z = (T)((x - mean) / sd) in double, 0 when sd < eps(T); out = (zr + zc) * (T)sqrt(1/2).
template <typename T>
__global__ void stouffer_tf_kernel(const T* __restrict__ X, int64_t ldx, int64_t M, int64_t N,
                                   const double* __restrict__ rmean, const double* __restrict__ rsd,
                                   const double* __restrict__ cmean, const double* __restrict__ csd, T* __restrict__ O,
                                   int64_t ldo) {
  const T isqrt2 = (T)0.70710678118654752440084436210485L;
  const int64_t total = M * N;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = t / N, j = t % N;
    const double x = (double)X[i * ldx + j];
    const T zr = rsd[i] < (double)Lim<T>::eps ? (T)0 : (T)((x - rmean[i]) / rsd[i]);
    const T zc = csd[j] < (double)Lim<T>::eps ? (T)0 : (T)((x - cmean[j]) / csd[j]);
    O[i * ldo + j] = (zr + zc) * isqrt2;
  }
}

int grid_of(int64_t total) { return (int)std::min<int64_t>(std::max<int64_t>(ceil_div(total, 256), 1), 65536); }

// Host staging helper: device copies of up to four M x N host matrices plus an output.
template <typename T>
struct Staged {
  const T* in[4] = {nullptr, nullptr, nullptr, nullptr};
  int64_t ld[4] = {0, 0, 0, 0};
  T* out = nullptr;
  int64_t ldo = 0;
};

template <typename T>
int combo_entry(mcpg_ctx* h, int ptrs, const T* MI, int64_t ldmi, const T* M1, int64_t ld1, const T* M2, int64_t ld2,
                const T* M3, int64_t ld3, int64_t M, int64_t N, const T* coeffs, T* out, int64_t ldo) {
  MCPG_CHECK_HANDLE(h);
  mcpg::begin_call(h);
  if (M < 0 || N < 0) return mcpg::set_error(h, MCPG_ERR_SHAPE, "negative dimension");
  if (M == 0 || N == 0) return MCPG_OK;
  if (!coeffs || !MI || !out) return mcpg::set_error(h, MCPG_ERR_INVALID_ARG, "null pointer");
  const double cf[4] = {(double)coeffs[0], (double)coeffs[1], (double)coeffs[2], (double)coeffs[3]};
  const T* in[4] = {MI, M1, M2, M3};
  const int64_t ldin[4] = {ldmi, ld1, ld2, ld3};
  for (int q = 0; q < 4; ++q) {
    const bool used = q == 0 || cf[q] != 0.0;
    if (used && !in[q]) return mcpg::set_error(h, MCPG_ERR_INVALID_ARG, "null matrix for a nonzero coefficient");
    if (used && ldin[q] < N) return mcpg::set_error(h, MCPG_ERR_INVALID_ARG, "leading dimension too small");
  }
  if (ldo < N) return mcpg::set_error(h, MCPG_ERR_INVALID_ARG, "leading dimension too small");
  if (ptrs != MCPG_HOST && ptrs != MCPG_DEVICE) return mcpg::set_error(h, MCPG_ERR_INVALID_ARG, "bad ptrs kind");
  MCPG_HIP(h, hipSetDevice(h->device));
  Staged<T> s;
  int st;
  if (ptrs == MCPG_HOST) {
    mcpg::Sizer sz;
    for (int q = 0; q < 4; ++q)
      if (q == 0 || cf[q] != 0.0) sz.add<T>(M * N);
    sz.add<T>(M * N);
    void* base = mcpg::workspace(h, sz.off, &st);
    if (!base) return st;
    mcpg::Carver cv(base);
    for (int q = 0; q < 4; ++q) {
      if (!(q == 0 || cf[q] != 0.0)) continue;
      T* d = cv.take<T>(M * N);
      if ((st = mcpg::copy_2d(h, d, N * sizeof(T), in[q], ldin[q] * sizeof(T), N * sizeof(T), M, hipMemcpyHostToDevice)))
        return st;
      s.in[q] = d;
      s.ld[q] = N;
    }
    s.out = cv.take<T>(M * N);
    s.ldo = N;
  } else {
    for (int q = 0; q < 4; ++q) { s.in[q] = in[q]; s.ld[q] = ldin[q]; }
    s.out = out;
    s.ldo = ldo;
  }
  {
    mcpg::LaunchScope ls(h, "combo_kernel");
    hipLaunchKernelGGL(combo_kernel<T>, dim3(grid_of(M * N)), dim3(256), 0, h->stream, s.in[0], s.ld[0], s.in[1], s.ld[1],
                       s.in[2], s.ld[2], s.in[3], s.ld[3], M, N, cf[0], cf[1], cf[2], cf[3], s.out, s.ldo);
  }
  if (ptrs == MCPG_HOST &&
      (st = mcpg::copy_2d(h, out, ldo * sizeof(T), s.out, N * sizeof(T), N * sizeof(T), M, hipMemcpyDeviceToHost)))
    return st;
  return mcpg::finish(h, "mcpg_combo");
}

template <typename T>
int stouffer_entry(mcpg_ctx* h, int ptrs, const T* X, int64_t ldx, int64_t N, int exclude_ones, T* out, int64_t ldo) {
  MCPG_CHECK_HANDLE(h);
  mcpg::begin_call(h);
  if (N < 0) return mcpg::set_error(h, MCPG_ERR_SHAPE, "negative dimension");
  if (N == 0) return MCPG_OK;
  if (!X || !out) return mcpg::set_error(h, MCPG_ERR_INVALID_ARG, "null matrix pointer");
  if (ldx < N || ldo < N) return mcpg::set_error(h, MCPG_ERR_INVALID_ARG, "leading dimension too small");
  if (ptrs != MCPG_HOST && ptrs != MCPG_DEVICE) return mcpg::set_error(h, MCPG_ERR_INVALID_ARG, "bad ptrs kind");
  if (N > (int64_t)INT32_MAX) return mcpg::set_error(h, MCPG_ERR_UNSUPPORTED, "matrix too large");
  MCPG_HIP(h, hipSetDevice(h->device));
  const bool host = ptrs == MCPG_HOST;
  mcpg::Sizer sz;
  if (host) { sz.add<T>(N * N); sz.add<T>(N * N); }
  sz.add<T>(N);
  sz.add<T>(N);
  int st;
  void* base = mcpg::workspace(h, sz.off, &st);
  if (!base) return st;
  mcpg::Carver cv(base);
  const T* dX = X;
  int64_t lx = ldx;
  T* dO = out;
  int64_t lo = ldo;
  if (host) {
    T* x = cv.take<T>(N * N);
    if ((st = mcpg::copy_2d(h, x, N * sizeof(T), X, ldx * sizeof(T), N * sizeof(T), N, hipMemcpyHostToDevice))) return st;
    dX = x; lx = N;
    dO = cv.take<T>(N * N); lo = N;
  }
  T* mean = cv.take<T>(N);
  T* sd = cv.take<T>(N);
  {
    mcpg::LaunchScope ls(h, "row_stats_kernel");
    hipLaunchKernelGGL(row_stats_kernel<T>, dim3((unsigned)N), dim3(256), 0, h->stream, dX, lx, N, exclude_ones, mean, sd);
  }
  {
    mcpg::LaunchScope ls(h, "stouffer_kernel");
    hipLaunchKernelGGL(stouffer_kernel<T>, dim3(grid_of(N * N)), dim3(256), 0, h->stream, dX, lx, N, N, mean, sd, dO, lo);
  }
  if (host && (st = mcpg::copy_2d(h, out, ldo * sizeof(T), dO, N * sizeof(T), N * sizeof(T), N, hipMemcpyDeviceToHost)))
    return st;
  return mcpg::finish(h, "mcpg_stouffer");
}

template <typename T>
int stouffer_tf_entry(mcpg_ctx* h, int ptrs, const T* X, int64_t ldx, int64_t M, int64_t N, T* out, int64_t ldo) {
  MCPG_CHECK_HANDLE(h);
  mcpg::begin_call(h);
  if (M < 0 || N < 0) return mcpg::set_error(h, MCPG_ERR_SHAPE, "negative dimension");
  if (M == 0 || N == 0) return MCPG_OK;
  if (!X || !out) return mcpg::set_error(h, MCPG_ERR_INVALID_ARG, "null matrix pointer");
  if (ldx < N || ldo < N) return mcpg::set_error(h, MCPG_ERR_INVALID_ARG, "leading dimension too small");
  if (ptrs != MCPG_HOST && ptrs != MCPG_DEVICE) return mcpg::set_error(h, MCPG_ERR_INVALID_ARG, "bad ptrs kind");
  if (M > (int64_t)INT32_MAX || ceil_div(N, 64) > (int64_t)INT32_MAX)
    return mcpg::set_error(h, MCPG_ERR_UNSUPPORTED, "matrix too large");
  MCPG_HIP(h, hipSetDevice(h->device));
  const bool host = ptrs == MCPG_HOST;
  mcpg::Sizer sz;
  if (host) { sz.add<T>(M * N); sz.add<T>(M * N); }
  sz.add<double>(2 * M);
  sz.add<double>(2 * N);
  int st;
  void* base = mcpg::workspace(h, sz.off, &st);
  if (!base) return st;
  mcpg::Carver cv(base);
  const T* dX = X;
  int64_t lx = ldx;
  T* dO = out;
  int64_t lo = ldo;
  if (host) {
    T* x = cv.take<T>(M * N);
    if ((st = mcpg::copy_2d(h, x, N * sizeof(T), X, ldx * sizeof(T), N * sizeof(T), M, hipMemcpyHostToDevice))) return st;
    dX = x; lx = N;
    dO = cv.take<T>(M * N); lo = N;
  }
  double* rm = cv.take<double>(2 * M);
  double* rs = rm + M;
  double* cm = cv.take<double>(2 * N);
  double* cs = cm + N;
  {
    mcpg::LaunchScope ls(h, "zrow_moments_kernel");
    hipLaunchKernelGGL(zrow_moments_kernel<T>, dim3((unsigned)M), dim3(256), 0, h->stream, dX, lx, N, rm, rs);
  }
  {
    mcpg::LaunchScope ls(h, "zcol_moments_kernel");
    hipLaunchKernelGGL(zcol_moments_kernel<T>, dim3((unsigned)ceil_div(N, 64)), dim3(256), 0, h->stream, dX, lx, M, N,
                       cm, cs);
  }
  {
    mcpg::LaunchScope ls(h, "stouffer_tf_kernel");
    hipLaunchKernelGGL(stouffer_tf_kernel<T>, dim3(grid_of(M * N)), dim3(256), 0, h->stream, dX, lx, M, N, rm, rs, cm, cs,
                       dO, lo);
  }
  if (host && (st = mcpg::copy_2d(h, out, ldo * sizeof(T), dO, N * sizeof(T), N * sizeof(T), M, hipMemcpyDeviceToHost)))
    return st;
  return mcpg::finish(h, "mcpg_stouffer_tf");
}

}  // namespace

extern "C" {

int mcpg_combo_f32(mcpg_handle h, int ptrs, const float* MI, int64_t ldmi, const float* M1, int64_t ld1, const float* M2,
                   int64_t ld2, const float* M3, int64_t ld3, int64_t M, int64_t N, const float* coeffs, float* out,
                   int64_t ldo) {
  return combo_entry<float>(h, ptrs, MI, ldmi, M1, ld1, M2, ld2, M3, ld3, M, N, coeffs, out, ldo);
}
int mcpg_combo_f64(mcpg_handle h, int ptrs, const double* MI, int64_t ldmi, const double* M1, int64_t ld1,
                   const double* M2, int64_t ld2, const double* M3, int64_t ld3, int64_t M, int64_t N,
                   const double* coeffs, double* out, int64_t ldo) {
  return combo_entry<double>(h, ptrs, MI, ldmi, M1, ld1, M2, ld2, M3, ld3, M, N, coeffs, out, ldo);
}
int mcpg_stouffer_f32(mcpg_handle h, int ptrs, const float* X, int64_t ldx, int64_t N, int exclude_ones, float* out,
                      int64_t ldo) {
  return stouffer_entry<float>(h, ptrs, X, ldx, N, exclude_ones, out, ldo);
}
int mcpg_stouffer_f64(mcpg_handle h, int ptrs, const double* X, int64_t ldx, int64_t N, int exclude_ones, double* out,
                      int64_t ldo) {
  return stouffer_entry<double>(h, ptrs, X, ldx, N, exclude_ones, out, ldo);
}

int mcpg_stouffer_tf_f32(mcpg_handle h, int ptrs, const float* X, int64_t ldx, int64_t M, int64_t N, float* out,
                         int64_t ldo) {
  return stouffer_tf_entry<float>(h, ptrs, X, ldx, M, N, out, ldo);
}
int mcpg_stouffer_tf_f64(mcpg_handle h, int ptrs, const double* X, int64_t ldx, int64_t M, int64_t N, double* out,
                         int64_t ldo) {
  return stouffer_tf_entry<double>(h, ptrs, X, ldx, M, N, out, ldo);
}

}  // extern "C"
