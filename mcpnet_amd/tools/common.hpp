// Shared pieces of the mi / mcp / mcpnet tools: common options, random input, timers.
#pragma once

#include <chrono>
#include <cstdint>
#include <cstdio>
#include <fstream>
#include <limits>
#include <string>
#include <unordered_set>
#include <vector>

#include "cli.hpp"
#include "matrix_io.hpp"
#include "splash/ds/aligned_matrix.hpp"

namespace tool {

// Common options of the reference (README.md:116-124 and splash::io::common_parameters call sites
// at apps/correlation/mi.cpp:83-93, :300); spellings of the options beyond -i/-o/-t/-m are this
// build's, the reference's own spellings live in splash, which its tree does not include.
inline void common_options(cli::Parser& p) {
  p.option("-i,--input", "input matrix (.exp, .csv, .tsv or .h5; rows = genes)");
  p.option("-o,--output", "output file");
  p.option("-t,--threads", "host threads (accepted for compatibility; the work runs on the GPU)", true, "1");
  p.option("--single", "use float (default: double, as the reference)", false);
  p.option("-r,--random", "use a random input matrix instead of -i", false);
  p.option("--rseed", "random seed", true, "11");
  p.option("--rmin", "random minimum", true, "0.0");
  p.option("--rmax", "random maximum", true, "1.0");
  p.option("--num-vectors", "rows of the random matrix", true, "100");
  p.option("--vector-size", "columns of the random matrix", true, "1000");
  p.option("--h5-group", "HDF5 group of the input matrix", true, "array");
  p.option("--h5-gene-key", "HDF5 dataset of the row (gene) names", true, "axis1");
  p.option("--h5-samples-key", "HDF5 dataset of the column (sample) names", true, "axis0");
  p.option("--h5-matrix-key", "HDF5 dataset of the values", true, "block0_values");
}

inline mio::H5Keys h5_keys(const cli::Parser& p) {
  mio::H5Keys k;
  k.group = p.str("--h5-group");
  k.genes = p.str("--h5-gene-key");
  k.samples = p.str("--h5-samples-key");
  k.matrix = p.str("--h5-matrix-key");
  return k;
}

inline uint64_t splitmix64(uint64_t seed, uint64_t ctr) {
  uint64_t z = seed + (ctr + 1) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

template <typename T>
splash::ds::aligned_matrix<T> random_matrix(uint64_t seed, double mn, double mx, size_t rows, size_t cols,
                                            std::vector<std::string>& rn, std::vector<std::string>& cn) {
  splash::ds::aligned_matrix<T> m(rows, cols);
  for (size_t i = 0; i < rows; ++i)
    for (size_t j = 0; j < cols; ++j) {
      const double u = (double)(splitmix64(seed, i * cols + j) >> 11) * 0x1.0p-53;
      m.at(i, j) = static_cast<T>(mn + (mx - mn) * u);
    }
  rn.clear();
  cn.clear();
  for (size_t i = 0; i < rows; ++i) rn.push_back("g" + std::to_string(i));
  for (size_t j = 0; j < cols; ++j) cn.push_back("s" + std::to_string(j));
  return m;
}

template <typename T>
splash::ds::aligned_matrix<T> load_input(const cli::Parser& p, std::vector<std::string>& genes,
                                         std::vector<std::string>& samples) {
  if (p.seen("--random"))
    return random_matrix<T>(p.integer("--rseed"), p.real("--rmin"), p.real("--rmax"), p.integer("--num-vectors"),
                            p.integer("--vector-size"), genes, samples);
  if (p.str("--input").empty()) throw std::runtime_error("no input: give -i FILE or --random");
  return mio::read_matrix<T>(p.str("--input"), genes, samples, h5_keys(p));
}

// Transcription factors: tfs[i] = max for a gene named in the file (one name per line, trimmed),
// lowest otherwise; tf_names lists the TF genes in gene order.  The limits are those of L:
// apps/filter/mcp.cpp:250-257 and dpi.cpp:180-187 take them from DataType, mcpnet_helper.hpp's
// load_tfs (:1074-1120) from double (so +-inf in a float vector).
template <typename T, typename L = T>
void load_tfs(const std::string& path, const std::vector<std::string>& genes, std::vector<T>& tfs,
              std::vector<std::string>& tf_names) {
  std::unordered_set<std::string> TFs;
  std::ifstream ifs(path);
  if (!ifs) throw std::runtime_error("cannot open " + path);
  std::string line;
  while (std::getline(ifs, line)) TFs.insert(mio::trim(line));
  tfs.clear();
  tf_names.clear();
  for (const auto& g : genes) {
    const bool is_tf = TFs.count(mio::trim(g)) > 0;
    tfs.push_back(static_cast<T>(is_tf ? std::numeric_limits<L>::max() : std::numeric_limits<L>::lowest()));
    if (is_tf) tf_names.push_back(g);
  }
  printf("TFs specified %zu, found %zu\n", TFs.size(), tf_names.size());
}

// The TF rows of a genes x columns matrix, in gene order (filter_mat_rows_by_tfs,
// mcpnet_helper.hpp:1126-1145).
template <typename T>
splash::ds::aligned_matrix<T> filter_rows_by_tfs(const splash::ds::aligned_matrix<T>& in, const std::vector<T>& tfs) {
  size_t n = 0;
  for (T v : tfs) n += v > (T)0;
  splash::ds::aligned_matrix<T> out(n, in.columns());
  for (size_t i = 0, j = 0; i < in.rows(); ++i)
    if (tfs[i] > (T)0) {
      std::copy(in.data(i), in.data(i) + in.columns(), out.data(j));
      ++j;
    }
  return out;
}

struct Timer {
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
  double seconds() const {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
};

}  // namespace tool
