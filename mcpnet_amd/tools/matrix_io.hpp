// Matrix files for the command-line tools.
//   .exp  the reference's EXP text format (data/r10c10.exp): a header "Id<TAB>Alias<TAB>col..."
//         then one row per gene "name<TAB>alias<TAB>values..."
//   .csv  comma separated, header ",col..." then "name,values..."
//   .tsv  tab separated, same layout as .csv
//   .h5   HDF5, pandas fixed layout (/array/axis0, axis1, block0_values), through h5io.hpp; a
//         build without the HDF5 header (MIO_NO_HDF5) rejects .h5 paths with a clear message
#pragma once

#include <cstdio>
#include <fstream>
#include <limits>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "splash/ds/aligned_matrix.hpp"
#ifndef MIO_NO_HDF5
#include "h5io.hpp"
#endif

namespace mio {

#ifdef MIO_NO_HDF5
struct H5Keys {
  std::string group = "array", genes = "axis1", samples = "axis0", matrix = "block0_values";
};
#endif

inline bool is_h5(const std::string& e) { return e == ".h5" || e == ".hdf5"; }

inline std::string ext(const std::string& p) {
  auto d = p.find_last_of('.');
  return d == std::string::npos ? "" : p.substr(d);
}

inline std::vector<std::string> split(const std::string& line, char sep) {
  std::vector<std::string> out;
  std::string cur;
  std::stringstream ss(line);
  while (std::getline(ss, cur, sep)) {
    if (!cur.empty() && cur.back() == '\r') cur.pop_back();
    out.push_back(cur);
  }
  return out;
}

inline std::string trim(const std::string& s) {
  auto b = s.find_first_not_of(" \t\r\n");
  auto e = s.find_last_not_of(" \t\r\n");
  return b == std::string::npos ? "" : s.substr(b, e - b + 1);
}

// rows = genes (first column names), columns = samples (header names)
template <typename T>
splash::ds::aligned_matrix<T> read_matrix(const std::string& path, std::vector<std::string>& rows,
                                          std::vector<std::string>& cols, const H5Keys& keys = H5Keys()) {
  const std::string e = ext(path);
#ifdef MIO_NO_HDF5
  if (is_h5(e)) throw std::runtime_error(path + ": this build has no HDF5 support; use .exp, .csv or .tsv");
#else
  if (is_h5(e)) return read_h5<T>(path, keys, rows, cols);
#endif
  const bool exp = e == ".exp";
  const char sep = (e == ".csv") ? ',' : '\t';
  const size_t skip = exp ? 2 : 1;
  std::ifstream in(path);
  if (!in) throw std::runtime_error("cannot open " + path);
  std::string line;
  if (!std::getline(in, line)) throw std::runtime_error(path + ": empty file");
  auto head = split(line, sep);
  if (head.size() < skip) throw std::runtime_error(path + ": malformed header");
  cols.assign(head.begin() + skip, head.end());
  std::vector<std::vector<T>> data;
  rows.clear();
  while (std::getline(in, line)) {
    if (trim(line).empty()) continue;
    auto f = split(line, sep);
    if (f.size() != cols.size() + skip) throw std::runtime_error(path + ": row '" + f[0] + "' has the wrong number of fields");
    rows.push_back(trim(f[0]));
    std::vector<T> v(cols.size());
    for (size_t j = 0; j < cols.size(); ++j) v[j] = static_cast<T>(std::stod(f[j + skip]));
    data.push_back(std::move(v));
  }
  splash::ds::aligned_matrix<T> m(data.size(), cols.size());
  for (size_t i = 0; i < data.size(); ++i)
    for (size_t j = 0; j < cols.size(); ++j) m.at(i, j) = data[i][j];
  return m;
}

template <typename T>
void write_matrix(const std::string& path, const std::vector<std::string>& rows, const std::vector<std::string>& cols,
                  const splash::ds::aligned_matrix<T>& m) {
  const std::string e = ext(path);
#ifdef MIO_NO_HDF5
  if (is_h5(e)) throw std::runtime_error(path + ": this build has no HDF5 support; use .exp, .csv or .tsv");
#else
  if (is_h5(e)) return write_h5<T>(path, "array", rows, cols, m);   // convert.cpp:72 writes group "array"
#endif
  const bool exp = e == ".exp";
  const char sep = (e == ".csv") ? ',' : '\t';
  std::ofstream out(path);
  if (!out) throw std::runtime_error("cannot write " + path);
  out << (exp ? "Id\tAlias" : "");
  for (size_t j = 0; j < m.columns(); ++j) out << (j || exp ? std::string(1, sep) : std::string(1, sep)) << (j < cols.size() ? cols[j] : std::to_string(j));
  out << "\n";
  char buf[64];
  for (size_t i = 0; i < m.rows(); ++i) {
    out << (i < rows.size() ? rows[i] : std::to_string(i));
    if (exp) out << "\t---";
    for (size_t j = 0; j < m.columns(); ++j) {
      snprintf(buf, sizeof buf, "%.*g", std::numeric_limits<T>::max_digits10, static_cast<double>(m.at(i, j)));
      out << sep << buf;
    }
    out << "\n";
  }
}

}  // namespace mio
