// Adaptive-partitioning MI: the reference's kernel functor names
// (include/mcp/correlation/AdaptivePartitioningMI.hpp:40-647 of MCPNet) with the patterns that
// apply them specialised to the GPU:
//   add_noise<T, Distribution, Generator>::operator()(in, out)          -> mcpg_add_noise_f32
//   Transform<M<IT>, IntegralCumulativeHistogramKernel<IT,0>, M<IT>>    -> mcpg_cumhist_i32
//   InnerProduct<M<IT>, AdaptivePartitionRankMIKernel2<IT,OT,0>, M<OT>> -> mcpg_ap_mi_i32 / _f64 (OT double)
// The reference's add_noise draws from per-thread std::default_random_engine streams of splash
// (absent); here the jitter is a seeded counter-based stream, identical for any thread count.
// Like the reference header, it brings in splash/kernel/random.hpp for the
// random_number_generator its callers construct.
#pragma once

#include <cstdint>
#include <random>
#include <type_traits>
#include <vector>

#include "mcpgpu/convert.hpp"
#include "splash/ds/aligned_matrix.hpp"
#include "splash/kernel/random.hpp"
#include "splash/transform/rank.hpp"
#include "splash/patterns/pattern.hpp"

namespace mcp {
namespace kernel {

// add_noise (AdaptivePartitioningMI.hpp:37-114): out = in + U[mn, mx), with the reference's
// template parameters and constructor (mcpnet_helper.hpp:111-114, mi.cpp:128-131).  The jitter
// is the GPU's counter-based stream under gen.seed(); only the range the reference's callers
// use, [0, 1e-5), is supported, and any other range is an MCPG_ERR_UNSUPPORTED error.
template <typename IT, typename Distribution = std::uniform_real_distribution<IT>,
          typename Generator = std::default_random_engine>
class add_noise {
  static_assert(std::is_same<Distribution, std::uniform_real_distribution<IT>>::value,
                "the GPU jitter is uniform");
  uint64_t seed_ = splash::kernel::random_number_generator<Generator>::default_seed;
  IT mn_ = 0.0, mx_ = 0.00001;

 public:
  add_noise() {}
  add_noise(splash::kernel::random_number_generator<Generator>& gen, IT const& mn = 0.0, IT const& mx = 1.0)
      : seed_(gen.seed()), mn_(mn), mx_(mx) {}
  virtual ~add_noise() {}
  void copy_parameters(const add_noise& o) {
    seed_ = o.seed_;
    mn_ = o.mn_;
    mx_ = o.mx_;
  }
  uint64_t seed() const { return seed_; }
  void operator()(const splash::ds::aligned_matrix<IT>& in, splash::ds::aligned_matrix<IT>& out) const {
    if (!(mn_ == static_cast<IT>(0.0) && mx_ == static_cast<IT>(0.00001)))
      throw mcpgpu::error(MCPG_ERR_UNSUPPORTED, "add_noise: the GPU jitter is U[0, 1e-5)");
    const size_t G = in.rows(), S = in.columns();
    std::vector<float> x = mcpgpu::packed<float>(in), y(G * S);
    mcpgpu::check(mcpg_add_noise_f32(mcpgpu::thread_handle(), MCPG_HOST, x.data(), G, S, S, seed_, y.data(), S),
                  "add_noise");
    mcpgpu::unpack(y, G, S, out);
  }
};

// IntegralCumulativeHistogramKernel (AdaptivePartitioningMI.hpp:414-442)
template <typename IT, long firstRank = 0>
class IntegralCumulativeHistogramKernel {
  static_assert(firstRank == 0, "ranks start at 0");

 public:
  void copy_parameters(const IntegralCumulativeHistogramKernel&) {}
};

// AdaptivePartitionRankMIKernel2 (AdaptivePartitioningMI.hpp:445-647).  The cumulative
// histograms given to the constructor are not needed: the GPU rebuilds them from the ranks.
template <typename IT, typename OT, long firstRank = 0>
class AdaptivePartitionRankMIKernel2 {
  static_assert(firstRank == 0, "ranks start at 0");

 public:
  AdaptivePartitionRankMIKernel2(IT const* /*rowPtrs*/, size_t const& /*stride_bytes*/) {}
  AdaptivePartitionRankMIKernel2() {}
  void copy_parameters(const AdaptivePartitionRankMIKernel2&) {}
};

}  // namespace kernel
}  // namespace mcp

namespace mcpgpu {
// The output arithmetic of AdaptivePartitionRankMIKernel2<IT, OT>: its result mi/S + log2 S is a
// double (AdaptivePartitioningMI.hpp:644), kept for OT = double, rounded to float otherwise.
template <typename OT>
struct ap_value {
  typedef float type;
};
template <>
struct ap_value<double> {
  typedef double type;
};
inline int ap_mi(const int32_t* r, size_t G, size_t S, float* mi) {
  return mcpg_ap_mi_i32(thread_handle(), MCPG_HOST, r, G, S, S, mi, G);
}
inline int ap_mi(const int32_t* r, size_t G, size_t S, double* mi) {
  return mcpg_ap_mi_f64(thread_handle(), MCPG_HOST, r, G, S, S, mi, G);
}
}  // namespace mcpgpu

namespace splash {
namespace pattern {

template <typename IT>
class Transform<ds::aligned_matrix<IT>, mcp::kernel::IntegralCumulativeHistogramKernel<IT, 0>, ds::aligned_matrix<IT>> {
 public:
  void operator()(const ds::aligned_matrix<IT>& in, const mcp::kernel::IntegralCumulativeHistogramKernel<IT, 0>&,
                  ds::aligned_matrix<IT>& out) const {
    const size_t G = in.rows(), S = in.columns();
    std::vector<int32_t> r = mcpgpu::packed<int32_t>(in), hst(G * (S + 1));
    mcpgpu::check(mcpg_cumhist_i32(mcpgpu::thread_handle(), MCPG_HOST, r.data(), G, S, S, hst.data(), S + 1),
                  "IntegralCumulativeHistogramKernel");
    mcpgpu::unpack(hst, G, S + 1, out);
  }
};

template <typename IT, typename OT, bool SYM>
class InnerProduct<ds::aligned_matrix<IT>, mcp::kernel::AdaptivePartitionRankMIKernel2<IT, OT, 0>, ds::aligned_matrix<OT>,
                   SYM> {
 public:
  void operator()(const ds::aligned_matrix<IT>& first, const ds::aligned_matrix<IT>& second,
                  const mcp::kernel::AdaptivePartitionRankMIKernel2<IT, OT, 0>&, ds::aligned_matrix<OT>& out) const {
    if (&first != &second && first.data() != second.data())
      throw mcpgpu::error(MCPG_ERR_INVALID_ARG, "AdaptivePartitionRankMIKernel2: the GPU path pairs rows of one rank matrix");
    const size_t G = first.rows(), S = first.columns();
    std::vector<int32_t> r = mcpgpu::packed<int32_t>(first);
    typedef typename mcpgpu::ap_value<OT>::type V;
    std::vector<V> mi(G * G);
    mcpgpu::check(mcpgpu::ap_mi(r.data(), G, S, mi.data()), "AdaptivePartitionRankMIKernel2");
    mcpgpu::unpack(mi, G, G, out);
  }
};

}  // namespace pattern
}  // namespace splash
