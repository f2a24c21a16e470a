/*
 * CPU oracle for MCPNet's hot paths -- TEST INFRASTRUCTURE, NOT PART OF THE PRODUCT.
 *
 * A plain-C restatement of the reference algorithms, used only as the checker by tests/,
 * __graft_entry__.smoke() and the cpu_baseline leg of bench.py.  Nothing in the product path
 * (libmcpgpu.so, mcpnet_amd, the command-line tools) may call it.  Compiled without fast-math
 * and with FMA contraction off, so the FP32 parts follow the reference's source-level float
 * arithmetic and the FP64 parts are the "FP64 oracle" that GPU results are measured against.
 *
 * Parity pins (checked in tests/test_oracle.py; see each function for its reference lines):
 *   knots                closed form [0]*order, i/(bins-order+1), [1]*order
 *   basis (f32 / f64)    scipy.interpolate.BSpline basis elements; partition of unity
 *   weights / entropy    FP64 marginals vs closed-form p log p of the basis sums
 *   MI (f64)             symmetry, MI(x,x) = 2H - H(x,x), invariance to sample permutation and
 *                        to positive affine maps, independent numpy evaluation from scipy bases
 *   AP-MI                closed form log2(S) - 2 for identical ranks with S = 2^k >= 8,
 *                        sum of leaf counts = S, an independent recursive Python restatement
 *   max-min              the reference's own simd_maxmin.hpp (built into oracle/_ref), and
 *                        brute force
 *   ratio                closed-form branches of combine.hpp:214-220
 *   MI gradient (f64)    central finite differences of orc_bspline_mi_loss_f64
 *   combo                unit coefficient rows reduce to the MCP2/3/4 ratios (orc_ratio) and to 1
 *   Stouffer             numpy mean / std (ddof 1) z-scores
 *   Stouffer (TF)        numpy mean / std (ddof 0) z-scores of rows and columns; the ddof-0 z-scores
 *                        give numpy's Pearson correlation as dot / count, as the reference's
 *                        PearsonKernel computes it from the same StandardScore<T, T, false>
 *   DPI                  an independent vectorised numpy statement; the weakest edge of a
 *                        triangle is removed at tolerance 0 (ARACNE's data processing inequality)
 */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#define MAXB 80
#define MAXO 8

/* ------------------------------------------------------------------ B-spline */

/* buildKnotVector<KT> (include/mcp/correlation/BSplineMI.hpp:35-71), KT = float. */
void orc_knots_f32(int bins, int order, float* v) {
  int n_int = bins - order;
  double norm = 1.0 / (double)(n_int + 1);
  for (int i = 0; i < order; ++i) v[i] = 0.0f;
  for (int i = order; i < order + n_int; ++i) v[i] = (float)((double)(float)(i - order + 1) * norm);
  for (int i = order + n_int; i < 2 * order + n_int; ++i) v[i] = 1.0f;
}

void orc_knots_f64(int bins, int order, double* v) {
  int n_int = bins - order;
  double norm = 1.0 / (double)(n_int + 1);
  for (int i = 0; i < order; ++i) v[i] = 0.0;
  for (int i = order; i < order + n_int; ++i) v[i] = (double)(i - order + 1) * norm;
  for (int i = order + n_int; i < 2 * order + n_int; ++i) v[i] = 1.0;
}

/* BSplineWeightsKernel::basisFunction (BSplineMI.hpp:130-164), recursive, float. */
float orc_basis_f32(const float* k, int bins, int i, int p, float t) {
  float d1, n1, d2, n2, e1, e2;
  if (p == 1) {
    if ((t >= k[i] && t < k[i + 1] && k[i] < k[i + 1]) || ((double)fabsf(t - k[i + 1]) < 1e-10 && (i + 1 == bins)))
      return 1.0f;
    return 0.0f;
  }
  d1 = k[i + p - 1] - k[i];
  n1 = t - k[i];
  d2 = k[i + p] - k[i + 1];
  n2 = k[i + p] - t;
  if ((double)d1 < 1e-10 && (double)d2 < 1e-10) {
    return 0.0f;
  } else if ((double)d1 < 1e-10) {
    e1 = 0.0f;
    e2 = n2 / d2 * orc_basis_f32(k, bins, i + 1, p - 1, t);
  } else if ((double)d2 < 1e-10) {
    e2 = 0.0f;
    e1 = n1 / d1 * orc_basis_f32(k, bins, i, p - 1, t);
  } else {
    e1 = n1 / d1 * orc_basis_f32(k, bins, i, p - 1, t);
    e2 = n2 / d2 * orc_basis_f32(k, bins, i + 1, p - 1, t);
  }
  if (e1 + e2 < 0) return 0.0f;
  return e1 + e2;
}

/* Same recursion in double (the FP64 oracle's basis). */
double orc_basis_f64(const double* k, int bins, int i, int p, double t) {
  double d1, n1, d2, n2, e1, e2;
  if (p == 1) {
    if ((t >= k[i] && t < k[i + 1] && k[i] < k[i + 1]) || (fabs(t - k[i + 1]) < 1e-10 && (i + 1 == bins))) return 1.0;
    return 0.0;
  }
  d1 = k[i + p - 1] - k[i];
  n1 = t - k[i];
  d2 = k[i + p] - k[i + 1];
  n2 = k[i + p] - t;
  if (d1 < 1e-10 && d2 < 1e-10) return 0.0;
  if (d1 < 1e-10) {
    e1 = 0.0;
    e2 = n2 / d2 * orc_basis_f64(k, bins, i + 1, p - 1, t);
  } else if (d2 < 1e-10) {
    e2 = 0.0;
    e1 = n1 / d1 * orc_basis_f64(k, bins, i, p - 1, t);
  } else {
    e1 = n1 / d1 * orc_basis_f64(k, bins, i, p - 1, t);
    e2 = n2 / d2 * orc_basis_f64(k, bins, i + 1, p - 1, t);
  }
  if (e1 + e2 < 0) return 0.0;
  return e1 + e2;
}

/* MinMaxScale<float> (BSplineMI.hpp:81-115): scale = 1.0/(max-min) in double stored as float;
 * a constant row keeps its values (scale 1, offset 0). */
void orc_minmax_scale_f32(const float* X, int64_t G, int64_t S, float* Y) {
  for (int64_t g = 0; g < G; ++g) {
    const float* x = X + g * S;
    float mn = S > 0 ? x[0] : 0.0f, mx = mn;
    for (int64_t j = 1; j < S; ++j) {
      mn = x[j] < mn ? x[j] : mn;
      mx = x[j] > mx ? x[j] : mx;
    }
    float scale = (mx > mn) ? (float)(1.0 / (double)(mx - mn)) : 1.0f;
    float offset = (mx > mn) ? mn : 0.0f;
    for (int64_t j = 0; j < S; ++j) Y[g * S + j] = scale * (x[j] - offset);
  }
}

/* BSplineWeightsKernel<float>::operator() (BSplineMI.hpp:248-273) on scaled rows: bin-major
 * float weights W[g][b*S+s] and, at W[g][bins*S], the marginal entropy.  The entropy follows
 * the FP64 marginal used by the GPU path: p_b = (sum_s w) / S in double, H = -sum p log2 p. */
void orc_bspline_weights_f32(const float* Xs, int64_t G, int64_t S, int bins, int order, float* W) {
  float k[MAXB + MAXO];
  orc_knots_f32(bins, order, k);
  const int64_t row = (int64_t)bins * S + 1;
#pragma omp parallel for schedule(static)
  for (int64_t g = 0; g < G; ++g) {
    float* w = W + g * row;
    for (int b = 0; b < bins; ++b)
      for (int64_t s = 0; s < S; ++s) w[b * S + s] = orc_basis_f32(k, bins, b, order, Xs[g * S + s]);
    double H = 0.0;
    for (int b = 0; b < bins; ++b) {
      double acc = 0.0;
      for (int64_t s = 0; s < S; ++s) acc += (double)w[b * S + s];
      double p = acc * (1.0 / (double)S);
      if (p > 0.0) H += p * log2(p);
    }
    w[bins * S] = (float)(0.0 - H);
  }
}

/* FP64 weights of raw rows: x scaled to [0,1] in double (constant rows kept), dense
 * W[g][b][s] and marginal entropy H[g]. */
/* Rows of float (Xf) or double (Xd) values; the scaling and everything after it in double. */
static void weights_f64_rows(const float* Xf, const double* Xd, int64_t G, int64_t S, int bins, int order, int do_scale,
                             double* W, double* H) {
  double k[MAXB + MAXO];
  orc_knots_f64(bins, order, k);
#pragma omp parallel for schedule(static)
  for (int64_t g = 0; g < G; ++g) {
#define XV(j) (Xf ? (double)Xf[g * S + (j)] : Xd[g * S + (j)])
    double mn = S > 0 ? XV(0) : 0.0, mx = mn;
    for (int64_t j = 1; j < S; ++j) {
      mn = XV(j) < mn ? XV(j) : mn;
      mx = XV(j) > mx ? XV(j) : mx;
    }
    double scale = (do_scale && mx > mn) ? 1.0 / (mx - mn) : 1.0;
    double offset = (do_scale && mx > mn) ? mn : 0.0;
    double* w = W + g * bins * S;
    double Hg = 0.0;
    for (int b = 0; b < bins; ++b) {
      double acc = 0.0;
      for (int64_t s = 0; s < S; ++s) {
        double t = (XV(s) - offset) * scale;
        double v = orc_basis_f64(k, bins, b, order, t);
        w[b * S + s] = v;
        acc += v;
      }
      double p = acc / (double)S;
      if (p > 0.0) Hg += p * log2(p);
    }
    H[g] = -Hg;
#undef XV
  }
}

void orc_bspline_weights_f64(const float* X, int64_t G, int64_t S, int bins, int order, int do_scale, double* W,
                             double* H) {
  weights_f64_rows(X, NULL, G, S, bins, order, do_scale, W, H);
}

/* The same on double rows: MinMaxScale<double> + BSplineWeightsKernel<double>, the reference's
 * double DataType (BSplineMI.hpp:75-116, :119-274). */
void orc_bspline_weights_x64(const double* X, int64_t G, int64_t S, int bins, int order, int do_scale, double* W,
                             double* H) {
  weights_f64_rows(NULL, X, G, S, bins, order, do_scale, W, H);
}

/* Sparse form of FP64 weight rows: per sample the first bin with a non-zero weight (or 0) and
 * the `order` consecutive values from there. */
static void sparse_rows(const double* W, int64_t G, int64_t S, int bins, int order, int* F, double* V) {
#pragma omp parallel for schedule(static)
  for (int64_t g = 0; g < G; ++g) {
    for (int64_t s = 0; s < S; ++s) {
      int f = 0;
      while (f < bins && W[(g * bins + f) * S + s] == 0.0) ++f;
      if (f == bins) f = 0;
      F[g * S + s] = f;
      for (int a = 0; a < order; ++a)
        V[(g * S + s) * order + a] = (f + a < bins) ? W[(g * bins + f + a) * S + s] : 0.0;
    }
  }
}

static double pair_mi_sparse(const int* F, const double* V, const double* H, int64_t S, int bins, int order, int64_t i,
                             int64_t j) {
  double J[MAXB * MAXB];
  memset(J, 0, sizeof(double) * bins * bins);
  for (int64_t s = 0; s < S; ++s) {
    const int fi = F[i * S + s], fj = F[j * S + s];
    const double* vi = V + (i * S + s) * order;
    const double* vj = V + (j * S + s) * order;
    for (int a = 0; a < order && fi + a < bins; ++a)
      for (int b = 0; b < order && fj + b < bins; ++b) J[(fi + a) * bins + fj + b] += vi[a] * vj[b];
  }
  double Hij = 0.0;
  for (int e = 0; e < bins * bins; ++e) {
    double p = J[e] / (double)S;
    if (p > 0.0) Hij += p * log2(p);
  }
  return H[i] + H[j] + Hij;  /* H_i + H_j - H_ij with H_ij = -sum p log2 p */
}

/* FP64 oracle of the full compute_mi pipeline (mcpnet_helper.hpp:56-99): MI for all pairs. */
void orc_bspline_mi_f64(const float* X, int64_t G, int64_t S, int bins, int order, double* MI) {
  double* W = (double*)malloc(sizeof(double) * G * bins * S);
  double* H = (double*)malloc(sizeof(double) * G);
  int* F = (int*)malloc(sizeof(int) * G * S);
  double* V = (double*)malloc(sizeof(double) * G * S * order);
  orc_bspline_weights_f64(X, G, S, bins, order, 1, W, H);
  sparse_rows(W, G, S, bins, order, F, V);
#pragma omp parallel for schedule(dynamic, 4)
  for (int64_t i = 0; i < G; ++i)
    for (int64_t j = i; j < G; ++j) {
      double v = pair_mi_sparse(F, V, H, S, bins, order, i, j);
      MI[i * G + j] = v;
      MI[j * G + i] = v;
    }
  free(W); free(H); free(F); free(V);
}

/* compute_mi on double rows (the reference's double DataType): all pairs. */
void orc_bspline_mi_x64(const double* X, int64_t G, int64_t S, int bins, int order, double* MI) {
  double* W = (double*)malloc(sizeof(double) * G * bins * S);
  double* H = (double*)malloc(sizeof(double) * G);
  int* F = (int*)malloc(sizeof(int) * G * S);
  double* V = (double*)malloc(sizeof(double) * G * S * order);
  orc_bspline_weights_x64(X, G, S, bins, order, 1, W, H);
  sparse_rows(W, G, S, bins, order, F, V);
#pragma omp parallel for schedule(dynamic, 4)
  for (int64_t i = 0; i < G; ++i)
    for (int64_t j = i; j < G; ++j) {
      double v = pair_mi_sparse(F, V, H, S, bins, order, i, j);
      MI[i * G + j] = v;
      MI[j * G + i] = v;
    }
  free(W); free(H); free(F); free(V);
}

/* FP64 MI for a list of pairs (sampled evaluation at sizes where all pairs are too many). */
void orc_bspline_mi_pairs_f64(const float* X, int64_t G, int64_t S, int bins, int order, const int32_t* pi,
                              const int32_t* pj, int64_t n, double* out) {
  double* W = (double*)malloc(sizeof(double) * G * bins * S);
  double* H = (double*)malloc(sizeof(double) * G);
  int* F = (int*)malloc(sizeof(int) * G * S);
  double* V = (double*)malloc(sizeof(double) * G * S * order);
  orc_bspline_weights_f64(X, G, S, bins, order, 1, W, H);
  sparse_rows(W, G, S, bins, order, F, V);
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t e = 0; e < n; ++e) out[e] = pair_mi_sparse(F, V, H, S, bins, order, pi[e], pj[e]);
  free(W); free(H); free(F); free(V);
}

/* ---- FP64 loss and analytic gradient of L = sum_n m_n MI(pi_n, pj_n) w.r.t. raw X. ----
 * Chain: X -> min-max scaling (sub-gradient at the first arg-min / arg-max) -> B-spline basis
 * (derivative of the order-`order` basis from the order-1 level) -> marginal and joint
 * probabilities -> entropies. */
double orc_bspline_mi_loss_f64(const float* X, int64_t G, int64_t S, int bins, int order, const int32_t* pi,
                               const int32_t* pj, const float* m, int64_t n) {
  double* out = (double*)malloc(sizeof(double) * (n > 0 ? n : 1));
  orc_bspline_mi_pairs_f64(X, G, S, bins, order, pi, pj, n, out);
  double L = 0.0;
  for (int64_t e = 0; e < n; ++e) L += (double)m[e] * out[e];
  free(out);
  return L;
}

/* The same loss on a double-precision input (finite differences need unrounded steps). */
double orc_bspline_mi_loss_x64(const double* X, int64_t G, int64_t S, int bins, int order, const int32_t* pi,
                               const int32_t* pj, const float* m, int64_t n) {
  double k[MAXB + MAXO];
  orc_knots_f64(bins, order, k);
  double L = 0.0;
  double* wi = (double*)malloc(sizeof(double) * bins * S);
  double* wj = (double*)malloc(sizeof(double) * bins * S);
  for (int64_t e = 0; e < n; ++e) {
    double Hs[2];
    for (int side = 0; side < 2; ++side) {
      const double* x = X + (int64_t)(side ? pj[e] : pi[e]) * S;
      double* w = side ? wj : wi;
      double mn = x[0], mx = x[0];
      for (int64_t s = 1; s < S; ++s) { mn = x[s] < mn ? x[s] : mn; mx = x[s] > mx ? x[s] : mx; }
      double scale = mx > mn ? 1.0 / (mx - mn) : 1.0, off = mx > mn ? mn : 0.0;
      double H = 0.0;
      for (int b = 0; b < bins; ++b) {
        double acc = 0.0;
        for (int64_t s = 0; s < S; ++s) {
          w[b * S + s] = orc_basis_f64(k, bins, b, order, (x[s] - off) * scale);
          acc += w[b * S + s];
        }
        double p = acc / (double)S;
        if (p > 0.0) H += p * log2(p);
      }
      Hs[side] = -H;
    }
    double Hij = 0.0;
    for (int b = 0; b < bins; ++b)
      for (int c = 0; c < bins; ++c) {
        double acc = 0.0;
        for (int64_t s = 0; s < S; ++s) acc += wi[b * S + s] * wj[c * S + s];
        double p = acc / (double)S;
        if (p > 0.0) Hij += p * log2(p);
      }
    L += (double)m[e] * (Hs[0] + Hs[1] + Hij);
  }
  free(wi); free(wj);
  return L;
}

static double dbasis_f64(const double* k, int bins, int b, int p, double t) {
  if (p == 1) return 0.0;
  double d1 = k[b + p - 1] - k[b], d2 = k[b + p] - k[b + 1], v = 0.0;
  if (d1 >= 1e-10) v += orc_basis_f64(k, bins, b, p - 1, t) / d1;
  if (d2 >= 1e-10) v -= orc_basis_f64(k, bins, b + 1, p - 1, t) / d2;
  return (double)(p - 1) * v;
}

void orc_bspline_mi_grad_f64(const float* X, int64_t G, int64_t S, int bins, int order, const int32_t* pi,
                             const int32_t* pj, const float* m, int64_t n, double* dX) {
  double k[MAXB + MAXO];
  orc_knots_f64(bins, order, k);
  double* W = (double*)malloc(sizeof(double) * G * bins * S);
  double* H = (double*)malloc(sizeof(double) * G);
  orc_bspline_weights_f64(X, G, S, bins, order, 1, W, H);
  double* dW = (double*)calloc((size_t)G * bins * S, sizeof(double));
  /* joint terms */
  for (int64_t e = 0; e < n; ++e) {
    const int64_t i = pi[e], j = pj[e];
    const double* wi = W + i * bins * S;
    const double* wj = W + j * bins * S;
    for (int b = 0; b < bins; ++b)
      for (int c = 0; c < bins; ++c) {
        double acc = 0.0;
        for (int64_t s = 0; s < S; ++s) acc += wi[b * S + s] * wj[c * S + s];
        double p = acc / (double)S;
        if (p <= 0.0) continue;
        /* dMI/dp_bc = -dH_ij/dp_bc = log2 p + 1/ln2 */
        double gbc = (double)m[e] * (log2(p) + 1.0 / log(2.0)) / (double)S;
        for (int64_t s = 0; s < S; ++s) {
          dW[(i * bins + b) * S + s] += gbc * wj[c * S + s];
          dW[(j * bins + c) * S + s] += gbc * wi[b * S + s];
        }
      }
    /* marginal terms: dMI/dp_b(i) = dH_i/dp_b = -(log2 p + 1/ln2) */
    for (int side = 0; side < 2; ++side) {
      const int64_t g = side ? j : i;
      for (int b = 0; b < bins; ++b) {
        double acc = 0.0;
        for (int64_t s = 0; s < S; ++s) acc += W[(g * bins + b) * S + s];
        double p = acc / (double)S;
        if (p <= 0.0) continue;
        double gb = -(double)m[e] * (log2(p) + 1.0 / log(2.0)) / (double)S;
        for (int64_t s = 0; s < S; ++s) dW[(g * bins + b) * S + s] += gb;
      }
    }
  }
  /* through the basis and the scaling */
  for (int64_t g = 0; g < G; ++g) {
    const float* x = X + g * S;
    double mn = x[0], mx = x[0];
    int64_t amn = 0, amx = 0;
    for (int64_t s = 1; s < S; ++s) {
      if (x[s] < mn) { mn = x[s]; amn = s; }
      if (x[s] > mx) { mx = x[s]; amx = s; }
    }
    const int constant = !(mx > mn);
    const double scale = constant ? 1.0 : 1.0 / (mx - mn), off = constant ? 0.0 : mn;
    double slo = 0.0, shi = 0.0;
    for (int64_t s = 0; s < S; ++s) {
      const double t = ((double)x[s] - off) * scale;
      double gs = 0.0;
      for (int b = 0; b < bins; ++b) gs += dW[(g * bins + b) * S + s] * dbasis_f64(k, bins, b, order, t);
      dX[g * S + s] = gs * scale;
      slo += gs * (t - 1.0);
      shi += gs * t;
    }
    if (!constant) {
      dX[g * S + amn] += slo * scale;
      dX[g * S + amx] -= shi * scale;
    }
  }
  free(W); free(H); free(dW);
}

/* ------------------------------------------------------------------ adaptive partitioning */

/* IntegralCumulativeHistogramKernel<IT,0> (AdaptivePartitioningMI.hpp:422-439), with the
 * zero-initialised output the reference relies on (mi.cpp:161). */
void orc_cumhist_i32(const int32_t* R, int64_t G, int64_t S, int32_t* H) {
  for (int64_t g = 0; g < G; ++g) {
    int32_t* out = H + g * (S + 1);
    memset(out, 0, sizeof(int32_t) * (S + 1));
    for (int64_t i = 0; i < S; ++i) ++out[R[g * S + i] + 1];
    for (int64_t i = 2; i <= S; ++i) out[i] += out[i - 1];
  }
}

typedef struct { int64_t x1, y1, x2, y2, n; int whole; } owin;

/* _getNumPointsInWindow (AdaptivePartitioningMI.hpp:495-523).  tests (nullable) accumulates the
 * number of points the scan compares, the reference's unit of AP-MI work. */
static int64_t points_in(const int32_t* cX, const int32_t* cY, const int32_t* xg, const owin* w, int64_t* tests) {
  if (w->x1 >= w->x2 || w->y1 >= w->y2) return 0;
  int64_t nx = cX[w->x2] - cX[w->x1];
  int64_t ny = cY[w->y2] - cY[w->y1];
  if (nx == 0 || ny == 0) return 0;
  int64_t beg = cY[w->y1], end = beg + ny, n = 0;
  for (int64_t i = beg; i < end; ++i) n += (xg[i] >= w->x1) && (xg[i] < w->x2);
  if (tests) *tests += ny;
  return n;
}

/* AdaptivePartitionRankMIKernel2::operator() (AdaptivePartitioningMI.hpp:553-645) for the pair
 * (x = row gx, y = row gy) of rank matrix R (G x S) with cumulative histograms Hc (G x (S+1)).
 * Leaves (nxy, nx, ny) are recorded in summation order when cells != NULL. */
static double ap_pair_impl(const int32_t* R, const int32_t* Hc, int64_t S, int64_t gx, int64_t gy, int32_t* cells,
                           int32_t max_leaves, int32_t* nleaves, int64_t* tests) {
  const int32_t* cX = Hc + gx * (S + 1);
  const int32_t* cY = Hc + gy * (S + 1);
  const int32_t* xr = R + gx * S;
  const int32_t* yr = R + gy * S;
  int32_t* temp = (int32_t*)calloc(S + 1, sizeof(int32_t));
  int32_t* xg = (int32_t*)calloc(S + 1, sizeof(int32_t));
  /* _createSparseMatrix (:462-493) */
  for (int64_t i = 0; i < S; ++i) {
    int64_t y = yr[i];
    int64_t off = cY[y] + temp[y];
    temp[y] += 1;
    xg[off] = xr[i];
  }
  int64_t cap = 64, sp = 0;
  owin* st = (owin*)malloc(sizeof(owin) * cap);
  st[sp++] = (owin){0, 0, S, S, S, 1};
  double mi = 0.0;
  int32_t leaf = 0;
  while (sp > 0) {
    owin w = st[sp - 1];
    int64_t mx = (w.x1 + w.x2) >> 1, my = (w.y1 + w.y2) >> 1;
    owin q[4] = {{w.x1, w.y1, mx, my, 0, 0}, {mx, w.y1, w.x2, my, 0, 0}, {w.x1, my, mx, w.y2, 0, 0}, {mx, my, w.x2, w.y2, 0, 0}};
    for (int i = 0; i < 4; ++i) q[i].n = points_in(cX, cY, xg, &q[i], tests);
    double e = (double)w.n / 4.0, chi = 0.0, d;
    for (int i = 0; i < 4; ++i) {
      d = (double)q[i].n - e;
      chi += d * d;
    }
    chi /= e;
    if ((chi > 7.815) || w.whole) {
      --sp;
      for (int i = 0; i < 4; ++i) {
        const int splittable = ((q[i].x1 + 1 < q[i].x2) || (q[i].y1 + 1 < q[i].y2)) && (q[i].n > 2);
        if (splittable) {
          if (sp == cap) { cap *= 2; st = (owin*)realloc(st, sizeof(owin) * cap); }
          st[sp++] = q[i];
        } else if (q[i].n > 0) {
          int64_t nx = cX[q[i].x2] - cX[q[i].x1], ny = cY[q[i].y2] - cY[q[i].y1];
          if (cells && leaf < max_leaves) {
            cells[3 * leaf] = (int32_t)q[i].n; cells[3 * leaf + 1] = (int32_t)nx; cells[3 * leaf + 2] = (int32_t)ny;
          }
          ++leaf;
          mi += (double)q[i].n * log2((double)q[i].n / (double)(nx * ny));
        }
      }
    } else if (w.n > 0) {
      int64_t nx = cX[w.x2] - cX[w.x1], ny = cY[w.y2] - cY[w.y1];
      if (cells && leaf < max_leaves) {
        cells[3 * leaf] = (int32_t)w.n; cells[3 * leaf + 1] = (int32_t)nx; cells[3 * leaf + 2] = (int32_t)ny;
      }
      ++leaf;
      mi += (double)w.n * log2((double)w.n / (double)(nx * ny));
      --sp;
    } else {
      --sp;
    }
  }
  free(st); free(temp); free(xg);
  if (nleaves) *nleaves = leaf;
  return mi / (double)S + log2((double)S);
}

double orc_ap_mi_pair(const int32_t* R, const int32_t* Hc, int64_t S, int64_t gx, int64_t gy, int32_t* cells,
                      int32_t max_leaves, int32_t* nleaves) {
  return ap_pair_impl(R, Hc, S, gx, gy, cells, max_leaves, nleaves, NULL);
}

/* Point comparisons of the reference's quadrant scans (every _getNumPointsInWindow call that scans)
 * over the pairs orc_ap_mi_rows evaluates for rows [row0, row0 + nrows): the algorithmic work
 * of AP-MI. */
int64_t orc_ap_tests_rows(const int32_t* R, int64_t G, int64_t S, int64_t row0, int64_t nrows) {
  int32_t* Hc = (int32_t*)malloc(sizeof(int32_t) * G * (S + 1));
  orc_cumhist_i32(R, G, S, Hc);
  int64_t total = 0;
#pragma omp parallel for schedule(dynamic, 8) collapse(2) reduction(+ : total)
  for (int64_t r = 0; r < nrows; ++r)
    for (int64_t j = 0; j < G; ++j) {
      const int64_t i = row0 + r;
      int64_t t = 0;
      if (i <= j) ap_pair_impl(R, Hc, S, i, j, NULL, 0, NULL, &t);
      else ap_pair_impl(R, Hc, S, j, i, NULL, 0, NULL, &t);
      total += t;
    }
  free(Hc);
  return total;
}

/* InnerProduct<..., AdaptivePartitionRankMIKernel2, ..., true>: i <= j evaluated as (x=i, y=j),
 * mirrored, FP32 output (mi.cpp:170-177). */
void orc_ap_mi(const int32_t* R, int64_t G, int64_t S, float* MI) {
  int32_t* Hc = (int32_t*)malloc(sizeof(int32_t) * G * (S + 1));
  orc_cumhist_i32(R, G, S, Hc);
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t i = 0; i < G; ++i)
    for (int64_t j = i; j < G; ++j) {
      float v = (float)orc_ap_mi_pair(R, Hc, S, i, j, NULL, 0, NULL);
      MI[i * G + j] = v;
      MI[j * G + i] = v;
    }
  free(Hc);
}

/* Rows [row0, row0+nrows) of the all-pairs AP-MI matrix (the CPU baseline sample): out is
 * nrows x G, pairs evaluated as (x = row, y = column) like the symmetric pattern's upper triangle. */
void orc_ap_mi_rows(const int32_t* R, int64_t G, int64_t S, int64_t row0, int64_t nrows, float* out) {
  int32_t* Hc = (int32_t*)malloc(sizeof(int32_t) * G * (S + 1));
  orc_cumhist_i32(R, G, S, Hc);
#pragma omp parallel for schedule(dynamic, 8) collapse(2)
  for (int64_t r = 0; r < nrows; ++r)
    for (int64_t j = 0; j < G; ++j) {
      const int64_t i = row0 + r;
      out[r * G + j] = (float)(i <= j ? orc_ap_mi_pair(R, Hc, S, i, j, NULL, 0, NULL) : orc_ap_mi_pair(R, Hc, S, j, i, NULL, 0, NULL));
    }
  free(Hc);
}

/* Single pair evaluation with leaf recording; Hc built for the two rows. */
double orc_ap_mi_cells(const int32_t* R, int64_t G, int64_t S, int64_t gx, int64_t gy, int32_t* cells,
                       int32_t max_leaves, int32_t* nleaves) {
  int32_t* Hc = (int32_t*)malloc(sizeof(int32_t) * G * (S + 1));
  orc_cumhist_i32(R, G, S, Hc);
  double v = orc_ap_mi_pair(R, Hc, S, gx, gy, cells, max_leaves, nleaves);
  free(Hc);
  return v;
}

/* ------------------------------------------------------------------ max-min and ratio */

/* max_of_pairmin_scalar (include/mcp/filter/simd_maxmin.hpp:30-46) for every (i, j), row . row;
 * the float accumulator starts at (float)numeric_limits<double>::lowest() == -inf. */
void orc_maxmin_f32(const float* A, int64_t lda, const float* B, int64_t ldb, const float* mask, int64_t M, int64_t N,
                    int64_t K, float* C, int64_t ldc) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < M; ++i)
    for (int64_t j = 0; j < N; ++j) {
      float val = (float)(-DBL_MAX);
      const float* x = A + i * lda;
      const float* y = B + j * ldb;
      for (int64_t k = 0; k < K; ++k) {
        float mn = (y[k] < x[k]) ? y[k] : x[k];            /* std::min(x, y) */
        if (mask) mn = (mn < mask[k]) ? mn : mask[k];      /* std::min(mask, .) */
        val = (val < mn) ? mn : val;                       /* std::max(val, .) */
      }
      C[i * ldc + j] = val;
    }
}

/* Double version; accumulator starts at lowest<double>() == -DBL_MAX (simd_maxmin.hpp:306). */
void orc_maxmin_f64(const double* A, int64_t lda, const double* B, int64_t ldb, const double* mask, int64_t M,
                    int64_t N, int64_t K, double* C, int64_t ldc) {
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < M; ++i)
    for (int64_t j = 0; j < N; ++j) {
      double val = -DBL_MAX;
      const double* x = A + i * lda;
      const double* y = B + j * ldb;
      for (int64_t k = 0; k < K; ++k) {
        double mn = (y[k] < x[k]) ? y[k] : x[k];
        if (mask) mn = (mn < mask[k]) ? mn : mask[k];
        val = (val < mn) ? mn : val;
      }
      C[i * ldc + j] = val;
    }
}

/* ratio_kernel<T,T> (include/mcp/transform/combine.hpp:214-220). */
void orc_ratio_f32(const float* A, const float* Mm, int64_t n, float* out) {
  for (int64_t i = 0; i < n; ++i)
    out[i] = (fabsf(A[i]) < FLT_EPSILON) ? 0.0f
             : (fabsf(Mm[i]) < FLT_EPSILON) ? FLT_MAX
                                            : (float)((double)A[i] / (double)Mm[i]);
}

void orc_ratio_f64(const double* A, const double* Mm, int64_t n, double* out) {
  for (int64_t i = 0; i < n; ++i)
    out[i] = (fabs(A[i]) < DBL_EPSILON) ? 0.0 : (fabs(Mm[i]) < DBL_EPSILON) ? DBL_MAX : A[i] / Mm[i];
}

/* ------------------------------------------------------------------ ensemble and filters */

/* compute_combo (apps/pipeline/mcpnet_helper.hpp:509-571): s accumulated in double, a coefficient
 * of 1 adds the value unscaled and 0 skips the term; the quotient is a double division stored in T. */
#define ORC_COMBO(NAME, T, EPS, BIG)                                                              \
  void NAME(const T* MI, const T* M1, const T* M2, const T* M3, int64_t n, const T* cs, T* out) { \
    const double a = cs[0], b = cs[1], c = cs[2], d = cs[3];                                      \
    for (int64_t i = 0; i < n; ++i) {                                                             \
      double s = 0.0;                                                                             \
      if (a == 1.0) s += MI[i]; else if (a != 0.0) s += a * MI[i];                                \
      if (b == 1.0) s += M1[i]; else if (b != 0.0) s += b * M1[i];                                \
      if (c == 1.0) s += M2[i]; else if (c != 0.0) s += c * M2[i];                                \
      if (d == 1.0) s += M3[i]; else if (d != 0.0) s += d * M3[i];                                \
      out[i] = (fabs((double)MI[i]) < (double)EPS) ? (T)0                                         \
               : (fabs(s) < (double)EPS) ? BIG : (T)((double)MI[i] / s);                          \
    }                                                                                             \
  }
ORC_COMBO(orc_combo_f32, float, FLT_EPSILON, FLT_MAX)
ORC_COMBO(orc_combo_f64, double, DBL_EPSILON, DBL_MAX)

/* compute_stouffer (apps/pipeline/mcpnet_helper.hpp:574-584): row mean and unbiased standard
 * deviation (sequential FP64 sums; entries equal to 1 skipped when exclude_ones), then
 * stouffer_vector_kernel (include/mcp/transform/stouffer.hpp:56-80) in T. */
#define ORC_STOUFFER(NAME, T, EPS)                                                                \
  void NAME(const T* X, int64_t N, int exclude_ones, T* out) {                                    \
    T* mean = (T*)malloc(sizeof(T) * (size_t)N);                                                  \
    T* sd = (T*)malloc(sizeof(T) * (size_t)N);                                                    \
    for (int64_t r = 0; r < N; ++r) {                                                             \
      const T* x = X + r * N;                                                                     \
      double s = 0.0, c = 0.0, q = 0.0;                                                           \
      for (int64_t j = 0; j < N; ++j) {                                                           \
        if (exclude_ones && x[j] == (T)1) continue;                                               \
        s += (double)x[j];                                                                        \
        c += 1.0;                                                                                 \
      }                                                                                           \
      const double mu = c > 0.0 ? s / c : 0.0;                                                    \
      for (int64_t j = 0; j < N; ++j) {                                                           \
        if (exclude_ones && x[j] == (T)1) continue;                                               \
        q += ((double)x[j] - mu) * ((double)x[j] - mu);                                           \
      }                                                                                           \
      mean[r] = (T)mu;                                                                            \
      sd[r] = (T)sqrt(c > 1.0 ? q / (c - 1.0) : 0.0);                                             \
    }                                                                                             \
    const T isqrt2 = (T)0.70710678118654752440084436210485L;                                      \
    for (int64_t i = 0; i < N; ++i) {                                                             \
      const T denom = (fabs((double)sd[i]) < (double)EPS) ? (T)0 : (T)(1.0 / (double)sd[i]);      \
      for (int64_t j = 0; j < N; ++j) {                                                           \
        const T x = X[i * N + j];                                                                 \
        const T zi = (T)((x - mean[i]) * denom);                                                  \
        const T zj = (fabs((double)sd[j]) < (double)EPS) ? (T)0 : (T)((x - mean[j]) / sd[j]);     \
        out[i * N + j] = (zi + zj) * isqrt2;                                                      \
      }                                                                                           \
    }                                                                                             \
    free(mean);                                                                                   \
    free(sd);                                                                                     \
  }
ORC_STOUFFER(orc_stouffer_f32, float, FLT_EPSILON)
ORC_STOUFFER(orc_stouffer_f64, double, DBL_EPSILON)

/* compute_stouffer_tf (apps/pipeline/mcpnet_helper.hpp:587-638): StandardScore<T, T, false> of every
 * row and every column of the M x N TFs-to-genes matrix (mean and population standard deviation in
 * double; z = (T)((x - mean) / sd), 0 when sd < eps(T)), then zscored_stouffer_kernel
 * (include/mcp/transform/stouffer.hpp:84-99): out = (zr + zc) * isqrt2 in T. */
#define ORC_STOUFFER_TF(NAME, T, EPS)                                                             \
  void NAME(const T* X, int64_t M, int64_t N, T* out) {                                           \
    double* rm = (double*)malloc(sizeof(double) * (size_t)M);                                     \
    double* rs = (double*)malloc(sizeof(double) * (size_t)M);                                     \
    double* cm = (double*)malloc(sizeof(double) * (size_t)N);                                     \
    double* cs = (double*)malloc(sizeof(double) * (size_t)N);                                     \
    for (int64_t i = 0; i < M; ++i) {                                                             \
      double s = 0.0, q = 0.0;                                                                    \
      for (int64_t j = 0; j < N; ++j) s += (double)X[i * N + j];                                  \
      const double mu = s / (double)N;                                                            \
      for (int64_t j = 0; j < N; ++j) q += ((double)X[i * N + j] - mu) * ((double)X[i * N + j] - mu); \
      rm[i] = mu;                                                                                 \
      rs[i] = sqrt(q / (double)N);                                                                \
    }                                                                                             \
    for (int64_t j = 0; j < N; ++j) {                                                             \
      double s = 0.0, q = 0.0;                                                                    \
      for (int64_t i = 0; i < M; ++i) s += (double)X[i * N + j];                                  \
      const double mu = s / (double)M;                                                            \
      for (int64_t i = 0; i < M; ++i) q += ((double)X[i * N + j] - mu) * ((double)X[i * N + j] - mu); \
      cm[j] = mu;                                                                                 \
      cs[j] = sqrt(q / (double)M);                                                                \
    }                                                                                             \
    const T isqrt2 = (T)0.70710678118654752440084436210485L;                                      \
    for (int64_t i = 0; i < M; ++i)                                                               \
      for (int64_t j = 0; j < N; ++j) {                                                           \
        const double x = (double)X[i * N + j];                                                    \
        const T zr = rs[i] < (double)EPS ? (T)0 : (T)((x - rm[i]) / rs[i]);                       \
        const T zc = cs[j] < (double)EPS ? (T)0 : (T)((x - cm[j]) / cs[j]);                       \
        out[i * N + j] = (zr + zc) * isqrt2;                                                      \
      }                                                                                           \
    free(rm);                                                                                     \
    free(rs);                                                                                     \
    free(cm);                                                                                     \
    free(cs);                                                                                     \
  }
ORC_STOUFFER_TF(orc_stouffer_tf_f32, float, FLT_EPSILON)
ORC_STOUFFER_TF(orc_stouffer_tf_f64, double, DBL_EPSILON)

/* DPI mask (include/mcp/filter/dpi.hpp:56-66 dpi_kernel and :86-105 dpi_tf_kernel, both with
 * EQUALITY_EXCLUSION) applied as mask<T, true> (include/mcp/filter/threshold.hpp:46-53): the
 * reference's early-exit scan over y, pair by pair.  dpi_kernel keeps its multiplier in double,
 * dpi_tf_kernel in T.  Returns the number of entries that changed. */
#define ORC_DPI(NAME, T)                                                                          \
  int64_t NAME(const T* A, int64_t N, T tolerance, const T* tfs, T* out) {                        \
    const double tm_d = 1.0 / (1.0 - (double)tolerance);                                          \
    const T tm_t = (T)(1.0 / (1.0 - (double)tolerance));                                          \
    int64_t removed = 0;                                                                          \
    _Pragma("omp parallel for reduction(+ : removed) schedule(static)")                           \
    for (int64_t x = 0; x < N; ++x) {                                                             \
      const T* rx = A + x * N;                                                                    \
      for (int64_t z = 0; z < N; ++z) {                                                           \
        const T* rz = A + z * N;                                                                  \
        int del = 1;                                                                              \
        if (x != z) {                                                                             \
          int64_t y = 0;                                                                          \
          if (tfs) {                                                                              \
            const T ixz = rx[z] * tm_t;                                                           \
            const int tf_xz = (tfs[x] > 0) ^ (tfs[z] > 0);                                        \
            for (; y < N && (y == x || y == z || rx[y] < ixz || rz[y] < ixz ||                    \
                             (tfs[y] <= 0 && tf_xz)); ++y) {}                                     \
          } else {                                                                                \
            const T ixz = (T)((double)rx[z] * tm_d);                                              \
            for (; y < N && (y == x || y == z || rx[y] < ixz || rz[y] < ixz); ++y) {}             \
          }                                                                                       \
          del = y < N;                                                                            \
        }                                                                                         \
        const T o = del ? (T)0 : rx[z];                                                           \
        out[x * N + z] = o;                                                                       \
        removed += (o != rx[z]);                                                                  \
      }                                                                                           \
    }                                                                                             \
    return removed;                                                                               \
  }
ORC_DPI(orc_dpi_f32, float)
ORC_DPI(orc_dpi_f64, double)
