// Row reductions of splash used by the Stouffer and CLR transforms (splash is not part of the
// reference tree; the names follow the call sites apps/pipeline/mcpnet_helper.hpp:580 and
// apps/transform/transform.cpp:216-218, whose comment reads "set zscore degree of freedom to 1 to
// use unbiased variance").  GaussianParamsExclude1<IT, OT, SAMPLE> reduces a row to (mean, stdev);
// this build reads SAMPLE = true as the unbiased (n - 1) variance and "Exclude1" as leaving entries
// equal to 1 out of the moments.  The GPU specialisations of the patterns that use it implement
// exactly that; the class itself carries no data.
#pragma once

namespace splash {
namespace kernel {

template <typename IT, typename OT = IT, bool SAMPLE = true>
class GaussianParamsExclude1 {
 public:
  static constexpr bool sample = SAMPLE;
  void copy_parameters(const GaussianParamsExclude1&) {}
};

}  // namespace kernel
}  // namespace splash
