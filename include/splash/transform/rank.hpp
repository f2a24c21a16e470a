// splash::kernel::Rank<IT, OT, firstRank, ascending> (used at apps/correlation/mi.cpp:141-147 and
// apps/pipeline/mcpnet_helper.hpp:127-133 with <T, size_t, 0, false>, commented "descending").
// Rank 0 goes to the largest value; ties are broken by ascending column index.  The Transform
// specialisation ranks every row on the GPU (mcpg_rank_f32).
#pragma once

#include <cstdint>
#include <vector>

#include "mcpgpu/convert.hpp"
#include "splash/patterns/pattern.hpp"

namespace splash {
namespace kernel {

template <typename IT, typename OT, long firstRank = 0, bool ascending = false>
class Rank {
  static_assert(firstRank == 0 && !ascending, "the GPU path implements descending ranks starting at 0");

 public:
  void copy_parameters(const Rank&) {}
};

}  // namespace kernel

namespace pattern {

template <typename IT, typename OT>
class Transform<ds::aligned_matrix<IT>, kernel::Rank<IT, OT, 0, false>, ds::aligned_matrix<OT>> {
 public:
  void operator()(const ds::aligned_matrix<IT>& in, const kernel::Rank<IT, OT, 0, false>&,
                  ds::aligned_matrix<OT>& out) const {
    const size_t G = in.rows(), S = in.columns();
    std::vector<float> x = mcpgpu::packed<float>(in);
    std::vector<int32_t> r(G * S);
    mcpgpu::check(mcpg_rank_f32(mcpgpu::thread_handle(), MCPG_HOST, x.data(), G, S, S, r.data(), S), "Rank");
    mcpgpu::unpack(r, G, S, out);
  }
};

}  // namespace pattern
}  // namespace splash
