// Elementwise kernels of the reference used by the MCP pipeline
// (include/mcp/transform/combine.hpp:42-58, :210-222 of MCPNet), applied on the GPU:
//   BinaryOp / GlobalBinaryOp<M, M, ratio_kernel<T,T>, M>  -> mcpg_ratio_f32 / _f64
//   BinaryOp / GlobalBinaryOp<M, M, max_kernel<T>, M>      -> mcpg_elementwise_max_f32 / _f64
#pragma once

#include <type_traits>

#include "splash/ds/aligned_matrix.hpp"
#include "splash/patterns/pattern.hpp"

namespace mcp {
namespace kernel {

// out = max(in, aux)  (combine.hpp:42-58)
template <typename IT>
class max_kernel {
 public:
  void copy_parameters(const max_kernel&) {}
};

// out = |in| < eps ? 0 : |aux| < eps ? max : (OT)((double)in / (double)aux)  (combine.hpp:210-222)
template <typename IT, typename OT = IT>
class ratio_kernel {
  static_assert(std::is_same<IT, OT>::value, "the GPU ratio keeps the element type");

 public:
  void copy_parameters(const ratio_kernel&) {}
};

}  // namespace kernel
}  // namespace mcp

namespace mcpgpu {
inline int ratio_call(mcpg_handle h, const float* a, int64_t la, const float* m, int64_t lm, int64_t M, int64_t N,
                      float* o, int64_t lo) {
  return mcpg_ratio_f32(h, MCPG_HOST, a, la, m, lm, M, N, o, lo);
}
inline int ratio_call(mcpg_handle h, const double* a, int64_t la, const double* m, int64_t lm, int64_t M, int64_t N,
                      double* o, int64_t lo) {
  return mcpg_ratio_f64(h, MCPG_HOST, a, la, m, lm, M, N, o, lo);
}
inline int max_call(mcpg_handle h, const float* a, int64_t la, const float* m, int64_t lm, int64_t M, int64_t N,
                    float* o, int64_t lo) {
  return mcpg_elementwise_max_f32(h, MCPG_HOST, a, la, m, lm, M, N, o, lo);
}
inline int max_call(mcpg_handle h, const double* a, int64_t la, const double* m, int64_t lm, int64_t M, int64_t N,
                    double* o, int64_t lo) {
  return mcpg_elementwise_max_f64(h, MCPG_HOST, a, la, m, lm, M, N, o, lo);
}
}  // namespace mcpgpu

namespace splash {
namespace pattern {

template <typename T>
class GlobalBinaryOp<ds::aligned_matrix<T>, ds::aligned_matrix<T>, mcp::kernel::ratio_kernel<T, T>, ds::aligned_matrix<T>> {
 public:
  void operator()(const ds::aligned_matrix<T>& in1, const ds::aligned_matrix<T>& in2, const mcp::kernel::ratio_kernel<T, T>&,
                  ds::aligned_matrix<T>& out) const {
    if (in1.rows() != in2.rows() || in1.columns() != in2.columns())
      throw mcpgpu::error(MCPG_ERR_SHAPE, "ratio_kernel: operand shapes differ");
    out.resize(in1.rows(), in1.columns());
    mcpgpu::check(mcpgpu::ratio_call(mcpgpu::thread_handle(), in1.data(), in1.ld(), in2.data(), in2.ld(), in1.rows(),
                                     in1.columns(), out.data(), out.ld()),
                  "ratio_kernel");
  }
};

// BinaryOp(in1, part1, in2, part2, kernel, out, outpart): rows part1.offset.. of in1 against rows
// part2.offset.. of in2 (mcpnet_helper.hpp:182-183); a single process uses whole matrices.
template <typename T>
class BinaryOp<ds::aligned_matrix<T>, ds::aligned_matrix<T>, mcp::kernel::ratio_kernel<T, T>, ds::aligned_matrix<T>> {
 public:
  void operator()(const ds::aligned_matrix<T>& in1, const utils::partition<size_t>& p1, const ds::aligned_matrix<T>& in2,
                  const utils::partition<size_t>& p2, const mcp::kernel::ratio_kernel<T, T>&, ds::aligned_matrix<T>& out,
                  const utils::partition<size_t>& po) const {
    const size_t rows = p2.size, cols = in2.columns();
    if (p1.offset + rows > in1.rows() || in1.columns() != cols)
      throw mcpgpu::error(MCPG_ERR_SHAPE, "ratio_kernel: operand shapes differ");
    out.resize(po.offset + rows, cols);
    mcpgpu::check(mcpgpu::ratio_call(mcpgpu::thread_handle(), in1.data(p1.offset), in1.ld(), in2.data(p2.offset), in2.ld(),
                                     rows, cols, out.data(po.offset), out.ld()),
                  "ratio_kernel");
  }
};

template <typename T>
class GlobalBinaryOp<ds::aligned_matrix<T>, ds::aligned_matrix<T>, mcp::kernel::max_kernel<T>, ds::aligned_matrix<T>> {
 public:
  void operator()(const ds::aligned_matrix<T>& in1, const ds::aligned_matrix<T>& in2, const mcp::kernel::max_kernel<T>&,
                  ds::aligned_matrix<T>& out) const {
    if (in1.rows() != in2.rows() || in1.columns() != in2.columns())
      throw mcpgpu::error(MCPG_ERR_SHAPE, "max_kernel: operand shapes differ");
    out.resize(in1.rows(), in1.columns());
    mcpgpu::check(mcpgpu::max_call(mcpgpu::thread_handle(), in1.data(), in1.ld(), in2.data(), in2.ld(), in1.rows(),
                                   in1.columns(), out.data(), out.ld()),
                  "max_kernel");
  }
};
template <typename T>
class BinaryOp<ds::aligned_matrix<T>, ds::aligned_matrix<T>, mcp::kernel::max_kernel<T>, ds::aligned_matrix<T>>
    : public GlobalBinaryOp<ds::aligned_matrix<T>, ds::aligned_matrix<T>, mcp::kernel::max_kernel<T>, ds::aligned_matrix<T>> {};

}  // namespace pattern
}  // namespace splash
