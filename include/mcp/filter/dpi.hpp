// Data-processing-inequality kernels of the reference (include/mcp/filter/dpi.hpp:42-110 of
// MCPNet) and the InnerProduct that applies them, on the GPU:
//   InnerProduct<M, dpi_kernel<T>, aligned_matrix<bool>>     -> mcpg_dpi_f32 / _f64 (mask output)
//   InnerProduct<M, dpi_tf_kernel<T>, aligned_matrix<bool>>  -> the same with the TF list
// mask(x, z) = true (remove) when x == z or some y != x, z (a TF when exactly one of x, z is a TF)
// has row_x[y] >= i_xz and row_z[y] >= i_xz, i_xz = row_x[z] / (1 - tolerance).  Only the
// EQUALITY_EXCLUSION = true form, which every reference application uses, is provided.
#pragma once

#include <cstdint>
#include <type_traits>
#include <vector>

#include "splash/ds/aligned_matrix.hpp"
#include "splash/patterns/pattern.hpp"

namespace mcp {
namespace kernel {

template <typename IT, bool EQUALITY_EXCLUSION = true, bool SYMMETRY = true>
class dpi_kernel {
  static_assert(EQUALITY_EXCLUSION, "only the EQUALITY_EXCLUSION form is provided");

 protected:
  IT tol;

 public:
  dpi_kernel(double const& tolerance = 0.1) : tol(static_cast<IT>(tolerance)) {}
  void copy_parameters(const dpi_kernel& o) { tol = o.tol; }
  IT tolerance() const { return tol; }
};

template <typename IT, bool EQUALITY_EXCLUSION = true>
class dpi_tf_kernel {
  static_assert(EQUALITY_EXCLUSION, "only the EQUALITY_EXCLUSION form is provided");

 protected:
  IT tol;
  std::vector<IT> TFs;

 public:
  dpi_tf_kernel(std::vector<IT> const& _TFs = std::vector<IT>(), IT const& tolerance = 0.1) : tol(tolerance), TFs(_TFs) {}
  void copy_parameters(const dpi_tf_kernel& o) { tol = o.tol; TFs = o.TFs; }
  IT tolerance() const { return tol; }
  const std::vector<IT>& tfs() const { return TFs; }
};

}  // namespace kernel
}  // namespace mcp

namespace mcpgpu {
inline int dpi_call(mcpg_handle h, const float* a, int64_t la, int64_t N, float tol, const float* tfs, float* o,
                    int64_t lo, uint8_t* m, int64_t lm, int64_t* removed) {
  return mcpg_dpi_f32(h, MCPG_HOST, a, la, N, tol, tfs, o, lo, m, lm, removed);
}
inline int dpi_call(mcpg_handle h, const double* a, int64_t la, int64_t N, double tol, const double* tfs, double* o,
                    int64_t lo, uint8_t* m, int64_t lm, int64_t* removed) {
  return mcpg_dpi_f64(h, MCPG_HOST, a, la, N, tol, tfs, o, lo, m, lm, removed);
}

template <typename T>
void dpi_mask(const splash::ds::aligned_matrix<T>& in, T tol, const T* tfs, splash::ds::aligned_matrix<bool>& mask) {
  if (in.rows() != in.columns()) throw error(MCPG_ERR_SHAPE, "dpi: input matrix should be square");
  const size_t N = in.rows();
  std::vector<uint8_t> m(N * N);
  check(dpi_call(thread_handle(), in.data(), in.ld(), N, tol, tfs, nullptr, 0, m.data(), N, nullptr), "dpi_kernel");
  mask.resize(N, N);
  for (size_t i = 0; i < N; ++i)
    for (size_t j = 0; j < N; ++j) mask.at(i, j) = m[i * N + j] != 0;
}
}  // namespace mcpgpu

namespace splash {
namespace pattern {

template <typename T, bool SYM>
class InnerProduct<ds::aligned_matrix<T>, mcp::kernel::dpi_kernel<T, true, SYM>, ds::aligned_matrix<bool>, false> {
 public:
  void operator()(const ds::aligned_matrix<T>& first, const ds::aligned_matrix<T>& second,
                  const mcp::kernel::dpi_kernel<T, true, SYM>& k, ds::aligned_matrix<bool>& out) const {
    if (first.data() != second.data()) throw mcpgpu::error(MCPG_ERR_INVALID_ARG, "dpi: both operands are the MI matrix");
    mcpgpu::dpi_mask(first, k.tolerance(), (const T*)nullptr, out);
  }
};

template <typename T>
class InnerProduct<ds::aligned_matrix<T>, mcp::kernel::dpi_tf_kernel<T, true>, ds::aligned_matrix<bool>, false> {
 public:
  void operator()(const ds::aligned_matrix<T>& first, const ds::aligned_matrix<T>& second,
                  const mcp::kernel::dpi_tf_kernel<T, true>& k, ds::aligned_matrix<bool>& out) const {
    if (first.data() != second.data()) throw mcpgpu::error(MCPG_ERR_INVALID_ARG, "dpi: both operands are the MI matrix");
    if (k.tfs().size() < first.rows()) throw mcpgpu::error(MCPG_ERR_SHAPE, "dpi_tf_kernel: TF list too short");
    mcpgpu::dpi_mask(first, k.tolerance(), k.tfs().data(), out);
  }
};

}  // namespace pattern
}  // namespace splash
