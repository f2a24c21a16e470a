// Max-min ("maximum capacity path") kernel functor of the reference
// (include/mcp/filter/mcp.hpp:34-93 of MCPNet) and the InnerProduct that applies it, on the GPU:
//   InnerProduct<M, mcp2_maxmin_kernel<T, MASKED, true>, M, SYM>  -> mcpg_maxmin_f32 / _f64
// out(x, z) = max_k min(first[x][k], second[z][k]) (row . row), and with MASKED the term is also
// min'ed with TFs[k] (FLT_MAX for a transcription factor, lowest otherwise, apps/filter/mcp.cpp:255).
// The ALLOW_LOOPS = false kernels (loop exclusion) are used by no reference application and are
// not provided.
#pragma once

#include <type_traits>
#include <vector>

#include "mcp/transform/combine.hpp"
#include "splash/ds/aligned_matrix.hpp"
#include "splash/patterns/pattern.hpp"

namespace mcp {
namespace kernel {

template <typename IT, bool MASKED = false, bool ALLOW_LOOPS = false>
class mcp2_maxmin_kernel;

template <typename IT, bool MASKED>
class mcp2_maxmin_kernel<IT, MASKED, true> {
 protected:
  std::vector<IT> TFs;

 public:
  mcp2_maxmin_kernel(std::vector<IT> const& _TFs = std::vector<IT>()) : TFs(_TFs) {}
  virtual ~mcp2_maxmin_kernel() {}
  void copy_parameters(mcp2_maxmin_kernel const& other) { TFs = other.TFs; }
  const std::vector<IT>& mask() const { return TFs; }
};

}  // namespace kernel
}  // namespace mcp

namespace mcpgpu {
inline int maxmin_call(mcpg_handle h, int ptrs, const float* A, int64_t lda, const float* B, int64_t ldb,
                       const float* km, int64_t M, int64_t N, int64_t K, int sym, float* C, int64_t ldc) {
  return mcpg_maxmin_f32(h, ptrs, A, lda, B, ldb, km, M, N, K, sym, C, ldc);
}
inline int maxmin_call(mcpg_handle h, int ptrs, const double* A, int64_t lda, const double* B, int64_t ldb,
                       const double* km, int64_t M, int64_t N, int64_t K, int sym, double* C, int64_t ldc) {
  return mcpg_maxmin_f64(h, ptrs, A, lda, B, ldb, km, M, N, K, sym, C, ldc);
}
}  // namespace mcpgpu

namespace splash {
namespace pattern {

template <typename IT, bool MASKED, bool SYM>
class InnerProduct<ds::aligned_matrix<IT>, mcp::kernel::mcp2_maxmin_kernel<IT, MASKED, true>, ds::aligned_matrix<IT>,
                   SYM> {
  static_assert(std::is_same<IT, float>::value || std::is_same<IT, double>::value, "float or double matrices");

 public:
  void operator()(const ds::aligned_matrix<IT>& first, const ds::aligned_matrix<IT>& second,
                  const mcp::kernel::mcp2_maxmin_kernel<IT, MASKED, true>& k, ds::aligned_matrix<IT>& out) const {
    const size_t M = first.rows(), N = second.rows(), K = first.columns();
    if (second.columns() != K) throw mcpgpu::error(MCPG_ERR_SHAPE, "mcp2_maxmin_kernel: column counts differ");
    if (MASKED && k.mask().size() < K) throw mcpgpu::error(MCPG_ERR_SHAPE, "mcp2_maxmin_kernel: TF mask too short");
    const bool same = (&first == &second) || (first.data() == second.data());
    out.resize(M, N);
    mcpgpu::check(mcpgpu::maxmin_call(mcpgpu::thread_handle(), MCPG_HOST, first.data(), first.ld(), second.data(),
                                      second.ld(), MASKED ? k.mask().data() : nullptr, M, N, K, (SYM && same) ? 1 : 0,
                                      out.data(), out.ld()),
                  "mcp2_maxmin_kernel");
  }
};

}  // namespace pattern
}  // namespace splash
