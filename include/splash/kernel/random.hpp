// splash::kernel::random_number_generator<Generator>: the generator object the reference's
// callers default-construct and hand to add_noise (apps/correlation/mi.cpp:129-131,
// apps/pipeline/mcpnet_helper.hpp:111-114); add_noise draws per-thread engines from it with
// get_generator(thread_id) (include/mcp/correlation/AdaptivePartitioningMI.hpp:43, :89).
// splash's own class is not part of the reference tree.
//
// Here the object carries the seed of the GPU's counter-based jitter stream: element (g, s)
// of the matrix add_noise perturbs draws counter g*S+s of that seed, so the jitter is the same
// for any thread count or row partition.  get_generator() still returns a seeded engine of
// type Generator for host code that draws its own numbers.
#pragma once

#include <cstddef>
#include <cstdint>
#include <random>

namespace splash {
namespace kernel {

template <typename Generator = std::default_random_engine>
class random_number_generator {
 public:
  using generator_type = Generator;
  // the `--noise-seed` default of the command-line tools
  static constexpr uint64_t default_seed = 1296388676612ull;

  explicit random_number_generator(uint64_t seed = default_seed) : seed_(seed) {}

  uint64_t seed() const { return seed_; }
  void seed(uint64_t s) { seed_ = s; }

  // an engine for thread `thread_id`, seeded from (seed, thread_id)
  Generator get_generator(size_t thread_id) const {
    std::seed_seq sq{static_cast<uint32_t>(seed_), static_cast<uint32_t>(seed_ >> 32),
                     static_cast<uint32_t>(thread_id)};
    return Generator(sq);
  }

 private:
  uint64_t seed_;
};

template <typename Generator>
constexpr uint64_t random_number_generator<Generator>::default_seed;

}  // namespace kernel
}  // namespace splash
