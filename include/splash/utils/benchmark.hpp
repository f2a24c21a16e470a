// Wall-clock helpers the reference's pipeline code brackets every stage with
// (getSysTime / get_duration_s, e.g. apps/pipeline/mcpnet_helper.hpp:82-85, :110-119,
// apps/correlation/mi.cpp:79-99).  splash's own header is not part of the reference tree.
#pragma once

#include <chrono>

inline std::chrono::steady_clock::time_point getSysTime() { return std::chrono::steady_clock::now(); }

inline double get_duration_s(std::chrono::steady_clock::time_point const& start,
                             std::chrono::steady_clock::time_point const& end) {
  return std::chrono::duration<double>(end - start).count();
}

inline double get_duration_ms(std::chrono::steady_clock::time_point const& start,
                              std::chrono::steady_clock::time_point const& end) {
  return std::chrono::duration<double, std::milli>(end - start).count();
}
