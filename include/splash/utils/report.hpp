// Progress-print macros of the reference's pipeline code (FMT_ROOT_PRINT, FMT_PRINT,
// FMT_PRINT_RT, FMT_ROOT_PRINT_RT, FMT_FLUSH; e.g. apps/pipeline/mcpnet_helper.hpp:84, :118,
// apps/correlation/mi.cpp:61-66).  splash's own header (fmt-based) is not part of the reference
// tree.  Format strings use "{}" placeholders with an optional ":[width][.precision][type]"
// spec (type f, e, g or d); "{{" and "}}" are literal braces.  One process is the whole job
// here, so the root-only forms print like the others.
#pragma once

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iomanip>
#include <sstream>
#include <string>

namespace splash {
namespace utils {
namespace detail {

template <typename A>
void fmt_arg(std::ostringstream& os, const char* spec, const char* end, const A& a) {
  std::ostringstream one;
  if (spec < end && *spec == ':') {
    ++spec;
    if (spec < end && (*spec == '<' || *spec == '>' || *spec == '^')) {
      if (*spec == '<') one << std::left;
      ++spec;
    }
    int width = 0;
    while (spec < end && *spec >= '0' && *spec <= '9') width = width * 10 + (*spec++ - '0');
    if (width) one << std::setw(width);
    if (spec < end && *spec == '.') {
      ++spec;
      int prec = 0;
      while (spec < end && *spec >= '0' && *spec <= '9') prec = prec * 10 + (*spec++ - '0');
      one << std::setprecision(prec);
    }
    if (spec < end) {
      if (*spec == 'f') one << std::fixed;
      else if (*spec == 'e') one << std::scientific;
    }
  }
  one << a;
  os << one.str();
}

inline void fmt_into(std::ostringstream& os, const char* f) {
  for (; *f; ++f) {
    if ((f[0] == '{' && f[1] == '{') || (f[0] == '}' && f[1] == '}')) ++f;
    os << *f;
  }
}

template <typename A, typename... Rest>
void fmt_into(std::ostringstream& os, const char* f, const A& a, const Rest&... rest) {
  for (; *f; ++f) {
    if ((f[0] == '{' && f[1] == '{') || (f[0] == '}' && f[1] == '}')) {
      os << *f++;
      continue;
    }
    if (f[0] == '{') {
      const char* close = std::strchr(f, '}');
      if (!close) break;
      fmt_arg(os, f + 1, close, a);
      fmt_into(os, close + 1, rest...);
      return;
    }
    os << *f;
  }
}

}  // namespace detail

template <typename... Args>
std::string format(const char* f, const Args&... args) {
  std::ostringstream os;
  detail::fmt_into(os, f, args...);
  return os.str();
}

}  // namespace utils
}  // namespace splash

#define FMT_PRINT(...) std::fputs(::splash::utils::format(__VA_ARGS__).c_str(), stdout)
#define FMT_PRINT_RT(...) FMT_PRINT(__VA_ARGS__)
#define FMT_ROOT_PRINT(...) FMT_PRINT(__VA_ARGS__)
#define FMT_ROOT_PRINT_RT(...) FMT_PRINT(__VA_ARGS__)
#define FMT_PRINT_MPI_ROOT(...) FMT_PRINT(__VA_ARGS__)
#define FMT_FLUSH() std::fflush(stdout)
