/*
 * CPU baseline port -- TEST INFRASTRUCTURE, NOT PART OF THE PRODUCT.
 *
 * The reference's OpenMP + SIMD FP32 B-spline MI inner product, restated so that it can be
 * timed on the host cores next to the GPU (bench.py's cpu_baseline leg).  Built with the
 * reference's release flags (-O2 -ffast-math -funroll-loops, CMakeLists.txt:21-25, plus
 * -fopenmp), so the compiler vectorises the dot products as it does for splash's dotp_*.
 */
#include <math.h>
#include <stdint.h>

/* splash::kernel::dotp_* stand-in: a plain FP32 dot product, vectorised under -ffast-math. */
static inline float dotp(const float* __restrict__ a, const float* __restrict__ b, int64_t n) {
  float s = 0.0f;
  for (int64_t i = 0; i < n; ++i) s += a[i] * b[i];
  return s;
}

/* BSplineEntropyKernel::entropy2d_simd + BSplineMIKernel::operator()
 * (include/mcp/correlation/BSplineMI.hpp:288-315, :492-499) on reference-layout weight rows
 * (G x (bins*S+1), entropy at [bins*S]); rows [row0, row0+nrows) against every gene. */
void port_bspline_mi_rows_f32(const float* W, int64_t G, int64_t S, int bins, int64_t row0, int64_t nrows, float* out) {
  const int64_t ld = (int64_t)bins * S + 1;
  const float norm = 1.0f / (float)S;
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t r = 0; r < nrows; ++r) {
    const float* wx = W + (row0 + r) * ld;
    for (int64_t j = 0; j < G; ++j) {
      const float* wy = W + j * ld;
      float H = 0.0f;
      for (int bx = 0; bx < bins; ++bx)
        for (int by = 0; by < bins; ++by) {
          float hx = dotp(wx + bx * S, wy + by * S, S);
          hx *= norm;
          if (hx > 0.0f) H += hx * log2f(hx);
        }
      out[r * G + j] = wx[bins * S] + wy[bins * S] - (0.0f - H);
    }
  }
}
