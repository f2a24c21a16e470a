// Stouffer combination of the reference (include/mcp/transform/stouffer.hpp:38-80 of MCPNet) and
// the row-reduce pattern that applies it, on the GPU:
//   ReduceTransform / GlobalReduceTransform<M, GaussianParamsExclude1<T,T,true>,
//                                           stouffer_vector_kernel<T>, M>   -> mcpg_stouffer_f32 / _f64
// out[i][j] = (z_i + z_j) / sqrt(2) with z_r = (in[i][j] - mean_r) / stdev_r (0 when stdev_r < eps),
// (mean_r, stdev_r) the Gaussian parameters of row r (compute_stouffer, apps/pipeline/mcpnet_helper.hpp:574-584).
#pragma once

#include <type_traits>

#include "splash/ds/aligned_matrix.hpp"
#include "splash/kernel/gaussian.hpp"
#include "splash/patterns/pattern.hpp"

namespace mcp {
namespace kernel {

template <typename IT, typename IT2 = IT, typename OT = IT>
class stouffer_vector_kernel {
  static_assert(std::is_same<IT, OT>::value, "the GPU Stouffer keeps the element type");

 public:
  void copy_parameters(const stouffer_vector_kernel&) {}
};

}  // namespace kernel
}  // namespace mcp

namespace mcpgpu {
inline int stouffer_call(mcpg_handle h, const float* x, int64_t lx, int64_t N, int excl, float* o, int64_t lo) {
  return mcpg_stouffer_f32(h, MCPG_HOST, x, lx, N, excl, o, lo);
}
inline int stouffer_call(mcpg_handle h, const double* x, int64_t lx, int64_t N, int excl, double* o, int64_t lo) {
  return mcpg_stouffer_f64(h, MCPG_HOST, x, lx, N, excl, o, lo);
}
// compute_stouffer_tf (apps/pipeline/mcpnet_helper.hpp:587-638): the StandardScore<T, T, false>
// normalisations of the TF rows and gene columns and the zscored_stouffer_kernel combination
// (stouffer.hpp:84-99) in one call.
inline int stouffer_tf_call(mcpg_handle h, const float* x, int64_t lx, int64_t M, int64_t N, float* o, int64_t lo) {
  return mcpg_stouffer_tf_f32(h, MCPG_HOST, x, lx, M, N, o, lo);
}
inline int stouffer_tf_call(mcpg_handle h, const double* x, int64_t lx, int64_t M, int64_t N, double* o, int64_t lo) {
  return mcpg_stouffer_tf_f64(h, MCPG_HOST, x, lx, M, N, o, lo);
}
}  // namespace mcpgpu

namespace splash {
namespace pattern {

template <typename T, typename IT2>
class ReduceTransform<ds::aligned_matrix<T>, kernel::GaussianParamsExclude1<T, T, true>,
                      mcp::kernel::stouffer_vector_kernel<T, IT2, T>, ds::aligned_matrix<T>> {
 public:
  void operator()(const ds::aligned_matrix<T>& in, const kernel::GaussianParamsExclude1<T, T, true>&,
                  const mcp::kernel::stouffer_vector_kernel<T, IT2, T>&, ds::aligned_matrix<T>& out) const {
    if (in.rows() != in.columns()) throw mcpgpu::error(MCPG_ERR_SHAPE, "stouffer: the score matrix must be square");
    out.resize(in.rows(), in.columns());
    mcpgpu::check(mcpgpu::stouffer_call(mcpgpu::thread_handle(), in.data(), in.ld(), in.rows(), 1, out.data(), out.ld()),
                  "stouffer_vector_kernel");
  }
};

template <typename T, typename IT2>
class GlobalReduceTransform<ds::aligned_matrix<T>, kernel::GaussianParamsExclude1<T, T, true>,
                            mcp::kernel::stouffer_vector_kernel<T, IT2, T>, ds::aligned_matrix<T>>
    : public ReduceTransform<ds::aligned_matrix<T>, kernel::GaussianParamsExclude1<T, T, true>,
                             mcp::kernel::stouffer_vector_kernel<T, IT2, T>, ds::aligned_matrix<T>> {};

}  // namespace pattern
}  // namespace splash
