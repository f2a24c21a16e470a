// B-spline mutual information: the reference's kernel functor names and parameters
// (include/mcp/correlation/BSplineMI.hpp:35-500 of MCPNet) with the splash patterns that apply
// them specialised to run on the GPU through the C ABI:
//   GlobalTransform<M, MinMaxScale<T>, M>            -> mcpg_minmax_scale_f32 / _f64
//   Transform<M, BSplineWeightsKernel<T>, M>         -> mcpg_bspline_weights_f32 / _f64
//   InnerProduct<M, BSplineMIKernel<T>, M, SYM>      -> mcpg_bspline_mi_weights_f32 / _f64
// The functors carry parameters only; there is no per-pair CPU evaluation.  DataType double runs
// the *_f64 entry points (scaling, weights and, at the default precision MCPG_FP64, the MI in
// double, as the reference's double DataType); a BSplineMIKernel<double> given another precision
// contracts FP32 images of the double weights and widens the MI.
#pragma once

#include <cassert>
#include <cstdint>
#include <type_traits>
#include <vector>

#include "mcpgpu/convert.hpp"
#include "splash/ds/aligned_matrix.hpp"
#include "splash/patterns/pattern.hpp"

namespace mcp {
namespace kernel {

// buildKnotVector (BSplineMI.hpp:35-71): `order` zeros, bins-order internal knots at
// i/(bins-order+1), `order` ones.
template <typename KT>
std::vector<KT> buildKnotVector(const int numBins, const int splineOrder) {
  std::vector<KT> v(numBins + splineOrder, 0.0);
  const int n_int = numBins - splineOrder;
  const double norm = 1.0 / static_cast<double>(n_int + 1);
  for (int i = splineOrder; i < splineOrder + n_int; ++i) v[i] = static_cast<KT>(i - splineOrder + 1) * norm;
  for (int i = splineOrder + n_int; i < 2 * splineOrder + n_int; ++i) v[i] = static_cast<KT>(1);
  return v;
}

// MinMaxScale (BSplineMI.hpp:75-116)
template <typename IT, typename OT = IT, bool SampleStats = true>
class MinMaxScale {
 public:
  using InputType = IT;
  using OutputType = OT;
  void copy_parameters(const MinMaxScale&) {}
};

// BSplineWeightsKernel (BSplineMI.hpp:119-274): rows of bins*S weights (bin-major) + entropy.
template <typename IT, typename OT = IT>
class BSplineWeightsKernel {
 public:
  int numBins = 10, splineOrder = 3, numSamples = 0;
  BSplineWeightsKernel(int const& bins, int const& order, int const& samples)
      : numBins(bins), splineOrder(order), numSamples(samples) {}
  BSplineWeightsKernel() {}
  void copy_parameters(const BSplineWeightsKernel& o) {
    numBins = o.numBins;
    splineOrder = o.splineOrder;
    numSamples = o.numSamples;
  }
  size_t get_num_bins() const { return numBins; }
};

// BSplineEntropyKernel / BSplineMIKernel (BSplineMI.hpp:276-336, :339-500).
template <typename IT, typename OT = IT>
class BSplineEntropyKernel {
 public:
  int numBins = 10, numSamples = 0;
  BSplineEntropyKernel(int const& bins, int const& samples) : numBins(bins), numSamples(samples) {}
  BSplineEntropyKernel() {}
  void copy_parameters(const BSplineEntropyKernel& o) {
    numBins = o.numBins;
    numSamples = o.numSamples;
  }
  virtual ~BSplineEntropyKernel() {}
};

template <typename IT, typename OT = IT>
class BSplineMIKernel : public BSplineEntropyKernel<IT, OT> {
 public:
  // contraction arithmetic of the GPU path: MCPG_FP32, MCPG_BF16, MCPG_BF16X3, MCPG_I8X3, or
  // MCPG_FP64 (double DataType only); by default the DataType's own
  static constexpr int kDefaultPrecision = std::is_same<OT, double>::value ? MCPG_FP64 : MCPG_FP32;
  int precision = kDefaultPrecision;
  BSplineMIKernel(int const& bins, int const& samples, int prec = kDefaultPrecision)
      : BSplineEntropyKernel<IT, OT>(bins, samples), precision(prec) {}
  BSplineMIKernel() {}
  void copy_parameters(const BSplineMIKernel& o) {
    BSplineEntropyKernel<IT, OT>::copy_parameters(o);
    precision = o.precision;
  }
};

}  // namespace kernel
}  // namespace mcp

namespace mcpgpu {
// The stage arithmetic of a DataType: double for double (the *_f64 entry points), else float.
template <typename T>
struct bspline_value {
  typedef float type;
};
template <>
struct bspline_value<double> {
  typedef double type;
};
inline int minmax_scale(const float* x, size_t G, size_t S, float* y) {
  return mcpg_minmax_scale_f32(thread_handle(), MCPG_HOST, x, G, S, S, y, S);
}
inline int minmax_scale(const double* x, size_t G, size_t S, double* y) {
  return mcpg_minmax_scale_f64(thread_handle(), MCPG_HOST, x, G, S, S, y, S);
}
inline int bspline_weights(const float* x, size_t G, size_t S, int bins, int order, float* w, size_t ldw) {
  return mcpg_bspline_weights_f32(thread_handle(), MCPG_HOST, x, G, S, S, bins, order, w, ldw);
}
inline int bspline_weights(const double* x, size_t G, size_t S, int bins, int order, double* w, size_t ldw) {
  return mcpg_bspline_weights_f64(thread_handle(), MCPG_HOST, x, G, S, S, bins, order, w, ldw);
}
}  // namespace mcpgpu

namespace splash {
namespace pattern {

template <typename T>
class GlobalTransform<ds::aligned_matrix<T>, mcp::kernel::MinMaxScale<T>, ds::aligned_matrix<T>> {
 public:
  void operator()(const ds::aligned_matrix<T>& in, const mcp::kernel::MinMaxScale<T>&, ds::aligned_matrix<T>& out) const {
    const size_t G = in.rows(), S = in.columns();
    typedef typename mcpgpu::bspline_value<T>::type V;
    std::vector<V> x = mcpgpu::packed<V>(in), y(G * S);
    mcpgpu::check(mcpgpu::minmax_scale(x.data(), G, S, y.data()), "MinMaxScale");
    mcpgpu::unpack(y, G, S, out);
  }
};
template <typename T>
class Transform<ds::aligned_matrix<T>, mcp::kernel::MinMaxScale<T>, ds::aligned_matrix<T>>
    : public GlobalTransform<ds::aligned_matrix<T>, mcp::kernel::MinMaxScale<T>, ds::aligned_matrix<T>> {};

template <typename T>
class Transform<ds::aligned_matrix<T>, mcp::kernel::BSplineWeightsKernel<T>, ds::aligned_matrix<T>> {
 public:
  void operator()(const ds::aligned_matrix<T>& in, const mcp::kernel::BSplineWeightsKernel<T>& k,
                  ds::aligned_matrix<T>& out) const {
    const size_t G = in.rows(), S = in.columns();
    assert(static_cast<int>(S) == k.numSamples && "sample count of the kernel and the matrix differ");
    const size_t row = static_cast<size_t>(k.numBins) * S + 1;
    typedef typename mcpgpu::bspline_value<T>::type V;
    std::vector<V> x = mcpgpu::packed<V>(in), w(G * row);
    mcpgpu::check(mcpgpu::bspline_weights(x.data(), G, S, k.numBins, k.splineOrder, w.data(), row),
                  "BSplineWeightsKernel");
    mcpgpu::unpack(w, G, row, out);
  }
};
template <typename T>
class GlobalTransform<ds::aligned_matrix<T>, mcp::kernel::BSplineWeightsKernel<T>, ds::aligned_matrix<T>>
    : public Transform<ds::aligned_matrix<T>, mcp::kernel::BSplineWeightsKernel<T>, ds::aligned_matrix<T>> {};

template <typename T, bool SYM>
class InnerProduct<ds::aligned_matrix<T>, mcp::kernel::BSplineMIKernel<T>, ds::aligned_matrix<T>, SYM> {
 public:
  // out(i, j) = MI of weight rows i and j; `first` and `second` are the same weight matrix in
  // every reference call site (mcpnet_helper.hpp:94, mi.cpp:251).
  void operator()(const ds::aligned_matrix<T>& first, const ds::aligned_matrix<T>& second,
                  const mcp::kernel::BSplineMIKernel<T>& k, ds::aligned_matrix<T>& out) const {
    if (&first != &second && first.data() != second.data())
      throw mcpgpu::error(MCPG_ERR_INVALID_ARG, "BSplineMIKernel: the GPU path pairs rows of one weight matrix");
    const size_t G = first.rows();
    const size_t S = static_cast<size_t>(k.numSamples);
    if (first.columns() != static_cast<size_t>(k.numBins) * S + 1)
      throw mcpgpu::error(MCPG_ERR_SHAPE, "BSplineMIKernel: weight rows must hold bins*samples+1 values");
    if (k.precision == MCPG_FP64) {
      if (!std::is_same<T, double>::value)
        throw mcpgpu::error(MCPG_ERR_INVALID_ARG, "BSplineMIKernel: MCPG_FP64 needs the double DataType");
      std::vector<double> w = mcpgpu::packed<double>(first), mi(G * G);
      mcpgpu::check(mcpg_bspline_mi_weights_f64(mcpgpu::thread_handle(), MCPG_HOST, w.data(), G, S, first.columns(),
                                                k.numBins, mi.data(), G),
                    "BSplineMIKernel");
      mcpgpu::unpack(mi, G, G, out);
      return;
    }
    std::vector<float> w = mcpgpu::packed<float>(first), mi(G * G);
    mcpgpu::check(mcpg_bspline_mi_weights_f32(mcpgpu::thread_handle(), MCPG_HOST, w.data(), G, S, first.columns(),
                                              k.numBins, k.precision, mi.data(), G),
                  "BSplineMIKernel");
    mcpgpu::unpack(mi, G, G, out);
  }
};

}  // namespace pattern
}  // namespace splash
