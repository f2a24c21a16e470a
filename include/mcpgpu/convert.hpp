// Host-side helpers of the pattern specialisations: contiguous float/double/int32 staging copies
// of aligned_matrix rows (the C ABI takes plain pointers with a leading dimension).
#pragma once

#include <cstdint>
#include <type_traits>
#include <vector>

#include "splash/ds/aligned_matrix.hpp"

namespace mcpgpu {

template <typename To, typename T>
std::vector<To> packed(const splash::ds::aligned_matrix<T>& m) {
  std::vector<To> v(m.rows() * m.columns());
  for (size_t i = 0; i < m.rows(); ++i)
    for (size_t j = 0; j < m.columns(); ++j) v[i * m.columns() + j] = static_cast<To>(m.at(i, j));
  return v;
}

template <typename From, typename T>
void unpack(const std::vector<From>& v, size_t rows, size_t cols, splash::ds::aligned_matrix<T>& m) {
  m.resize(rows, cols);
  for (size_t i = 0; i < rows; ++i)
    for (size_t j = 0; j < cols; ++j) m.at(i, j) = static_cast<T>(v[i * cols + j]);
}

}  // namespace mcpgpu
