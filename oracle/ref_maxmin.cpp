// Thin C entry points around the reference's own max_of_pairmin
// (include/mcp/filter/simd_maxmin.hpp:30-450) -- TEST INFRASTRUCTURE.
// The header is compiled from its place in the reference tree (it needs nothing beyond
// <x86intrin.h>); the result goes to oracle/_ref/ and is used as the max-min parity oracle and
// as the CPU baseline of the max-min product.  Loops mirror the InnerProduct pattern: one
// max_of_pairmin call per output element, OpenMP over output rows.
#include <cstddef>
#include <cstdint>

#include "mcp/filter/simd_maxmin.hpp"

extern "C" {

void ref_maxmin_f32(const float* A, int64_t lda, const float* B, int64_t ldb, const float* mask, int64_t M, int64_t N,
                    int64_t K, float* C, int64_t ldc) {
#pragma omp parallel for schedule(dynamic, 4)
  for (int64_t i = 0; i < M; ++i)
    for (int64_t j = 0; j < N; ++j)
      C[i * ldc + j] = static_cast<float>(static_cast<double>(
          mcp::kernel::max_of_pairmin(A + i * lda, B + j * ldb, static_cast<size_t>(K), mask)));
}

void ref_maxmin_f64(const double* A, int64_t lda, const double* B, int64_t ldb, const double* mask, int64_t M,
                    int64_t N, int64_t K, double* C, int64_t ldc) {
#pragma omp parallel for schedule(dynamic, 4)
  for (int64_t i = 0; i < M; ++i)
    for (int64_t j = 0; j < N; ++j)
      C[i * ldc + j] = mcp::kernel::max_of_pairmin(A + i * lda, B + j * ldb, static_cast<size_t>(K), mask);
}

}  // extern "C"
