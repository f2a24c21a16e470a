// Pairwise and row-wise pattern templates of splash (the reference's engine; its source is not
// part of the reference tree, so the shapes follow the call sites: InnerProduct at
// apps/pipeline/mcpnet_helper.hpp:90-98, :157-162, :193-195; Transform at :70-80, :127-133;
// GlobalTransform at :60-64; BinaryOp at :182-183; GlobalBinaryOp at apps/filter/mcp.cpp:493-500).
//
// Only specialisations exist: each one, declared next to its kernel functor in include/mcp/...,
// sends the whole matrix to the GPU through the C ABI in include/mcpgpu/mcpgpu.h.  There is no
// generic (CPU, per-element) implementation: instantiating a pattern with a functor that has no
// GPU specialisation is a compile-time error.
#pragma once

#include <stdexcept>
#include <string>
#include <vector>

#include "mcpgpu/mcpgpu.h"

namespace splash {
namespace pattern {

// out(x, z) = kernel(first.row(x), second.row(z)); SYMMETRIC = first and second are the same
// matrix and only one triangle is evaluated (then mirrored).
template <typename InMat, typename Kernel, typename OutMat, bool SYMMETRIC = false>
class InnerProduct;

// out.row(i) = kernel(in.row(i)) for every row.
template <typename InMat, typename Kernel, typename OutMat>
class Transform;

// Same as Transform; in the reference the output is row-distributed over MPI ranks.
template <typename InMat, typename Kernel, typename OutMat>
class GlobalTransform;

// out.row(i) = kernel(in1.row(i), in2.row(i)).
template <typename InMat1, typename InMat2, typename Kernel, typename OutMat>
class BinaryOp;

template <typename InMat1, typename InMat2, typename Kernel, typename OutMat>
class GlobalBinaryOp;

// params = reduc(in.row(r)) for every row r, then out.row(i) = kernel(in.row(i), &params[i], params)
// (ReduceTransform / GlobalReduceTransform at apps/pipeline/mcpnet_helper.hpp:574-584 and
// apps/transform/transform.cpp:218-235).
template <typename InMat, typename Reduc, typename Kernel, typename OutMat>
class ReduceTransform;

template <typename InMat, typename Reduc, typename Kernel, typename OutMat>
class GlobalReduceTransform;

}  // namespace pattern

namespace utils {

// row range of a matrix, as the reference's BinaryOp receives it (partition<size_t>(offset, size, id))
template <typename S>
struct partition {
  S offset = 0, size = 0, id = 0;
  partition() = default;
  partition(S o, S s, S i) : offset(o), size(s), id(i) {}
  static partition make_partition(S n) { return partition(0, n, 0); }
};

}  // namespace utils
}  // namespace splash

namespace mcpgpu {

// Error raised by the pattern specialisations (the C ABI returns status codes).
struct error : std::runtime_error {
  int status;
  error(int st, const std::string& what) : std::runtime_error(what), status(st) {}
};

// One libmcpgpu handle per host thread (the reference copies kernel functors per thread).
inline mcpg_handle thread_handle(int device = 0) {
  struct Holder {
    mcpg_handle h = nullptr;
    ~Holder() {
      if (h) mcpg_destroy(h);
    }
  };
  thread_local Holder holder;
  if (!holder.h) {
    int st = mcpg_create(&holder.h, device);
    if (st != MCPG_OK) {
      holder.h = nullptr;
      throw error(st, std::string("mcpg_create: ") + mcpg_status_string(st));
    }
  }
  return holder.h;
}

inline void check(int st, const char* where) {
  if (st != MCPG_OK) {
    mcpg_handle h = thread_handle();
    throw error(st, std::string(where) + ": " + mcpg_status_string(st) + " (" + mcpg_last_error(h) + ")");
  }
}

}  // namespace mcpgpu
