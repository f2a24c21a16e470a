// splash::ds::aligned_matrix<T> -- the host matrix type the reference's patterns and apps use
// (row-major, rows = genes, padded rows; interface as called in apps/pipeline/mcpnet_helper.hpp
// and apps/filter/mcp.cpp: rows(), columns(), data(), data(i), column_bytes(), at(i, j),
// resize(), zero(), allgather(), scatter(), local_transpose()).  Single process: allgather()
// and scatter() are the identity (the reference's MPI row-block distribution is removed).
#pragma once

#include <cstddef>
#include <cstdlib>
#include <cstring>
#include <new>
#include <utility>

namespace splash {
namespace ds {

template <typename T>
class aligned_matrix {
 public:
  using value_type = T;
  static constexpr size_t kAlign = 64;

  aligned_matrix() = default;
  aligned_matrix(size_t rows, size_t cols) { resize(rows, cols); }
  aligned_matrix(const aligned_matrix& o) { copy_from(o); }
  aligned_matrix(aligned_matrix&& o) noexcept { swap(o); }
  aligned_matrix& operator=(const aligned_matrix& o) {
    if (this != &o) copy_from(o);
    return *this;
  }
  aligned_matrix& operator=(aligned_matrix&& o) noexcept {
    swap(o);
    return *this;
  }
  ~aligned_matrix() { std::free(data_); }

  size_t rows() const { return rows_; }
  size_t columns() const { return cols_; }
  // padded row stride in bytes (a multiple of 64)
  size_t column_bytes() const { return ld_ * sizeof(T); }
  size_t ld() const { return ld_; }

  T* data() { return data_; }
  const T* data() const { return data_; }
  T* data(size_t i) { return data_ + i * ld_; }
  const T* data(size_t i) const { return data_ + i * ld_; }
  T& at(size_t i, size_t j) { return data_[i * ld_ + j]; }
  const T& at(size_t i, size_t j) const { return data_[i * ld_ + j]; }
  T& operator()(size_t i, size_t j) { return at(i, j); }
  const T& operator()(size_t i, size_t j) const { return at(i, j); }

  void resize(size_t rows, size_t cols) {
    if (rows == rows_ && cols == cols_ && data_) return;
    std::free(data_);
    data_ = nullptr;
    rows_ = rows;
    cols_ = cols;
    const size_t per = kAlign / sizeof(T) > 0 ? kAlign / sizeof(T) : 1;
    ld_ = (cols + per - 1) / per * per;
    if (ld_ == 0) ld_ = per;
    const size_t bytes = rows_ * ld_ * sizeof(T);
    if (bytes) {
      void* p = nullptr;
      if (posix_memalign(&p, kAlign, (bytes + kAlign - 1) / kAlign * kAlign) != 0 || !p) throw std::bad_alloc();
      data_ = static_cast<T*>(p);
      std::memset(data_, 0, bytes);
    }
  }
  void zero() {
    if (data_) std::memset(data_, 0, rows_ * ld_ * sizeof(T));
  }
  aligned_matrix allgather() const { return *this; }
  aligned_matrix& scatter(int /*root*/) { return *this; }
  aligned_matrix local_transpose() const {
    aligned_matrix t(cols_, rows_);
    for (size_t i = 0; i < rows_; ++i)
      for (size_t j = 0; j < cols_; ++j) t.at(j, i) = at(i, j);
    return t;
  }

 private:
  void copy_from(const aligned_matrix& o) {
    resize(o.rows_, o.cols_);
    if (data_ && o.data_) std::memcpy(data_, o.data_, rows_ * ld_ * sizeof(T));
  }
  void swap(aligned_matrix& o) noexcept {
    std::swap(data_, o.data_);
    std::swap(rows_, o.rows_);
    std::swap(cols_, o.cols_);
    std::swap(ld_, o.ld_);
  }
  T* data_ = nullptr;
  size_t rows_ = 0, cols_ = 0, ld_ = 0;
};

}  // namespace ds
}  // namespace splash
