// HDF5 matrices in the pandas "fixed" layout that the reference's data files use
// (data/r10c10.h5, data/r127c999.h5, data/r128c1000.h5) and that its tools write through
// splash's write_matrix(path, "array", genes, samples, m) (apps/transform/convert.cpp:72):
//
//   /<group>/axis0          column names (samples), fixed-length NUL-terminated strings
//   /<group>/axis1          row names (genes)
//   /<group>/block0_items   the column names again
//   /<group>/block0_values  rows x columns values (F64 or F32), row-major
//
// plus the PyTables/pandas attributes (CLASS, VERSION, pandas_type "frame", transposed, ...) so
// that pandas.read_hdf(path, "array") reads the file back.  The dataset names are options of the
// reader, as in the reference (common_params.h5_group / h5_gene_key / h5_samples_key /
// h5_matrix_key at apps/correlation/mi.cpp:90-93); the defaults are the names the data files use.
//
// libhdf5 is opened at run time (dlopen), so the tools start, and read and write text formats,
// on machines that have no HDF5 library; an .h5 path then fails with a message naming the library.
// The header is only needed at build time, for the C types of the functions.
#pragma once

#include <dlfcn.h>
#include <hdf5.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "splash/ds/aligned_matrix.hpp"

namespace mio {

struct H5Keys {
  std::string group = "array";
  std::string genes = "axis1";
  std::string samples = "axis0";
  std::string matrix = "block0_values";
};

namespace h5 {

// H5F_ACC_RDONLY / H5F_ACC_TRUNC (H5Fpublic.h); the header's macros call H5check and H5open, which
// this file only reaches through dlsym.
constexpr unsigned kAccRdonly = 0x0000u;
constexpr unsigned kAccTrunc = 0x0002u;

// The functions and predefined type ids this file uses, resolved from libhdf5 once per process.
struct Lib {
#define MIO_H5_FN(name) decltype(&::name) name = nullptr;
  MIO_H5_FN(H5open)
  MIO_H5_FN(H5Eset_auto2)
  MIO_H5_FN(H5Fopen)
  MIO_H5_FN(H5Fcreate)
  MIO_H5_FN(H5Fclose)
  MIO_H5_FN(H5Gcreate2)
  MIO_H5_FN(H5Gopen2)
  MIO_H5_FN(H5Gclose)
  MIO_H5_FN(H5Lexists)
  MIO_H5_FN(H5Dopen2)
  MIO_H5_FN(H5Dcreate2)
  MIO_H5_FN(H5Dread)
  MIO_H5_FN(H5Dwrite)
  MIO_H5_FN(H5Dget_space)
  MIO_H5_FN(H5Dget_type)
  MIO_H5_FN(H5Dclose)
  MIO_H5_FN(H5Dvlen_reclaim)
  MIO_H5_FN(H5Screate)
  MIO_H5_FN(H5Screate_simple)
  MIO_H5_FN(H5Sget_simple_extent_ndims)
  MIO_H5_FN(H5Sget_simple_extent_dims)
  MIO_H5_FN(H5Sclose)
  MIO_H5_FN(H5Tcopy)
  MIO_H5_FN(H5Tset_size)
  MIO_H5_FN(H5Tset_strpad)
  MIO_H5_FN(H5Tset_cset)
  MIO_H5_FN(H5Tget_class)
  MIO_H5_FN(H5Tget_size)
  MIO_H5_FN(H5Tis_variable_str)
  MIO_H5_FN(H5Tclose)
  MIO_H5_FN(H5Acreate2)
  MIO_H5_FN(H5Awrite)
  MIO_H5_FN(H5Aclose)
#undef MIO_H5_FN
  hid_t C_S1 = -1, NATIVE_DOUBLE = -1, NATIVE_FLOAT = -1, NATIVE_LLONG = -1, NATIVE_B8 = -1;
  hid_t IEEE_F64LE = -1, IEEE_F32LE = -1, STD_I64LE = -1, STD_B8LE = -1;
  std::string error;

  Lib() {
    void* so = nullptr;
    const char* env = std::getenv("MCPG_HDF5_LIB");
    for (const char* cand : {env, "libhdf5.so", "libhdf5.so.103", "/opt/conda/lib/libhdf5.so"}) {
      if (cand && (so = dlopen(cand, RTLD_NOW | RTLD_LOCAL))) break;
    }
    if (!so) {
      error = "the HDF5 library (libhdf5.so) could not be loaded; set MCPG_HDF5_LIB to its path";
      return;
    }
    auto sym = [&](const char* n) {
      void* p = dlsym(so, n);
      if (!p && error.empty()) error = std::string("libhdf5 lacks ") + n;
      return p;
    };
#define MIO_H5_FN(name) name = reinterpret_cast<decltype(name)>(sym(#name));
    MIO_H5_FN(H5open)
    MIO_H5_FN(H5Eset_auto2)
    MIO_H5_FN(H5Fopen)
    MIO_H5_FN(H5Fcreate)
    MIO_H5_FN(H5Fclose)
    MIO_H5_FN(H5Gcreate2)
    MIO_H5_FN(H5Gopen2)
    MIO_H5_FN(H5Gclose)
    MIO_H5_FN(H5Lexists)
    MIO_H5_FN(H5Dopen2)
    MIO_H5_FN(H5Dcreate2)
    MIO_H5_FN(H5Dread)
    MIO_H5_FN(H5Dwrite)
    MIO_H5_FN(H5Dget_space)
    MIO_H5_FN(H5Dget_type)
    MIO_H5_FN(H5Dclose)
    MIO_H5_FN(H5Dvlen_reclaim)
    MIO_H5_FN(H5Screate)
    MIO_H5_FN(H5Screate_simple)
    MIO_H5_FN(H5Sget_simple_extent_ndims)
    MIO_H5_FN(H5Sget_simple_extent_dims)
    MIO_H5_FN(H5Sclose)
    MIO_H5_FN(H5Tcopy)
    MIO_H5_FN(H5Tset_size)
    MIO_H5_FN(H5Tset_strpad)
    MIO_H5_FN(H5Tset_cset)
    MIO_H5_FN(H5Tget_class)
    MIO_H5_FN(H5Tget_size)
    MIO_H5_FN(H5Tis_variable_str)
    MIO_H5_FN(H5Tclose)
    MIO_H5_FN(H5Acreate2)
    MIO_H5_FN(H5Awrite)
    MIO_H5_FN(H5Aclose)
#undef MIO_H5_FN
    if (!error.empty()) return;
    if (H5open() < 0) {
      error = "H5open failed";
      return;
    }
    // the predefined types are global hid_t variables, valid after H5open
    auto tid = [&](const char* n) {
      const hid_t* p = static_cast<const hid_t*>(sym(n));
      return p ? *p : (hid_t)-1;
    };
    C_S1 = tid("H5T_C_S1_g");
    NATIVE_DOUBLE = tid("H5T_NATIVE_DOUBLE_g");
    NATIVE_FLOAT = tid("H5T_NATIVE_FLOAT_g");
    NATIVE_LLONG = tid("H5T_NATIVE_LLONG_g");
    NATIVE_B8 = tid("H5T_NATIVE_B8_g");
    IEEE_F64LE = tid("H5T_IEEE_F64LE_g");
    IEEE_F32LE = tid("H5T_IEEE_F32LE_g");
    STD_I64LE = tid("H5T_STD_I64LE_g");
    STD_B8LE = tid("H5T_STD_B8LE_g");
    if (!error.empty()) return;
    H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);   // errors are reported by the callers' exceptions
  }
};

inline Lib& lib() {
  static Lib L;
  if (!L.error.empty()) throw std::runtime_error("HDF5: " + L.error);
  return L;
}

// owns one hid_t and closes it with the matching H5*close
struct Id {
  hid_t id = -1;
  herr_t (*close)(hid_t) = nullptr;
  Id(hid_t i, herr_t (*c)(hid_t)) : id(i), close(c) {}
  Id(const Id&) = delete;
  Id& operator=(const Id&) = delete;
  ~Id() {
    if (id >= 0 && close) close(id);
  }
  operator hid_t() const { return id; }
};

inline hid_t check(hid_t id, const std::string& what) {
  if (id < 0) throw std::runtime_error("HDF5: " + what);
  return id;
}

inline void check(herr_t st, const std::string& what, int) {
  if (st < 0) throw std::runtime_error("HDF5: " + what);
}

inline std::string join(const std::string& group, const std::string& name) {
  std::string g = group;
  while (!g.empty() && g.back() == '/') g.pop_back();
  if (!g.empty() && g.front() != '/') g = "/" + g;
  return g + "/" + name;
}

inline bool exists(hid_t file, const std::string& path) {
  Lib& L = lib();
  // every component must exist before H5Lexists may be asked about the next
  size_t pos = 1;
  while (true) {
    const size_t nxt = path.find('/', pos);
    const std::string part = path.substr(0, nxt);
    if (part.size() > 1 && L.H5Lexists(file, part.c_str(), H5P_DEFAULT) <= 0) return false;
    if (nxt == std::string::npos) return true;
    pos = nxt + 1;
  }
}

// one-dimensional dataset of names: fixed or variable length strings, or integers
inline std::vector<std::string> read_names(hid_t file, const std::string& path) {
  Lib& L = lib();
  Id ds(check(L.H5Dopen2(file, path.c_str(), H5P_DEFAULT), "cannot open " + path), L.H5Dclose);
  Id sp(check(L.H5Dget_space(ds), "dataspace of " + path), L.H5Sclose);
  if (L.H5Sget_simple_extent_ndims(sp) != 1) throw std::runtime_error("HDF5: " + path + " is not one-dimensional");
  hsize_t n = 0;
  L.H5Sget_simple_extent_dims(sp, &n, nullptr);
  Id ft(check(L.H5Dget_type(ds), "type of " + path), L.H5Tclose);
  std::vector<std::string> out(n);
  const H5T_class_t cls = L.H5Tget_class(ft);
  if (cls == H5T_STRING && L.H5Tis_variable_str(ft) > 0) {
    Id mt(L.H5Tcopy(L.C_S1), L.H5Tclose);
    L.H5Tset_size(mt, H5T_VARIABLE);
    std::vector<char*> buf(n, nullptr);
    check(L.H5Dread(ds, mt, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf.data()), "reading " + path, 0);
    for (hsize_t i = 0; i < n; ++i) out[i] = buf[i] ? buf[i] : "";
    L.H5Dvlen_reclaim(mt, sp, H5P_DEFAULT, buf.data());
  } else if (cls == H5T_STRING) {
    const size_t w = L.H5Tget_size(ft);
    Id mt(L.H5Tcopy(L.C_S1), L.H5Tclose);
    L.H5Tset_size(mt, w);
    L.H5Tset_strpad(mt, H5T_STR_NULLPAD);
    std::vector<char> buf(n * w + 1, 0);
    check(L.H5Dread(ds, mt, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf.data()), "reading " + path, 0);
    for (hsize_t i = 0; i < n; ++i) {
      const char* s = buf.data() + i * w;
      out[i].assign(s, strnlen(s, w));
    }
  } else if (cls == H5T_INTEGER) {
    std::vector<long long> buf(n);
    check(L.H5Dread(ds, L.NATIVE_LLONG, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf.data()), "reading " + path, 0);
    for (hsize_t i = 0; i < n; ++i) out[i] = std::to_string(buf[i]);
  } else {
    throw std::runtime_error("HDF5: " + path + " holds neither strings nor integers");
  }
  return out;
}

template <typename T>
hid_t native() {
  return std::is_same<T, float>::value ? lib().NATIVE_FLOAT : lib().NATIVE_DOUBLE;
}

}  // namespace h5

// Reads rows x columns values and their names.  Missing name datasets give "0", "1", ...
template <typename T>
splash::ds::aligned_matrix<T> read_h5(const std::string& path, const H5Keys& keys, std::vector<std::string>& rows,
                                      std::vector<std::string>& cols) {
  using h5::Id;
  h5::Lib& L = h5::lib();
  Id f(h5::check(L.H5Fopen(path.c_str(), h5::kAccRdonly, H5P_DEFAULT), "cannot open " + path), L.H5Fclose);
  const std::string mpath = h5::join(keys.group, keys.matrix);
  if (!h5::exists(f, mpath)) throw std::runtime_error(path + ": no dataset " + mpath);
  Id ds(h5::check(L.H5Dopen2(f, mpath.c_str(), H5P_DEFAULT), "cannot open " + mpath), L.H5Dclose);
  Id sp(h5::check(L.H5Dget_space(ds), "dataspace of " + mpath), L.H5Sclose);
  if (L.H5Sget_simple_extent_ndims(sp) != 2) throw std::runtime_error(path + ": " + mpath + " is not a matrix");
  hsize_t dims[2] = {0, 0};
  L.H5Sget_simple_extent_dims(sp, dims, nullptr);
  const size_t R = dims[0], C = dims[1];
  splash::ds::aligned_matrix<T> m(R, C);
  if (R && C) {
    std::vector<T> buf(R * C);
    h5::check(L.H5Dread(ds, h5::native<T>(), H5S_ALL, H5S_ALL, H5P_DEFAULT, buf.data()), "reading " + mpath, 0);
    for (size_t i = 0; i < R; ++i) std::copy(buf.begin() + i * C, buf.begin() + (i + 1) * C, m.data(i));
  }
  auto names = [&](const std::string& key, size_t n) {
    const std::string p = h5::join(keys.group, key);
    std::vector<std::string> v;
    if (!key.empty() && h5::exists(f, p)) v = h5::read_names(f, p);
    else
      for (size_t i = 0; i < n; ++i) v.push_back(std::to_string(i));
    return v;
  };
  rows = names(keys.genes, R);
  cols = names(keys.samples, C);
  return m;
}

template <typename T>
void write_h5(const std::string& path, const std::string& group, const std::vector<std::string>& rows,
              const std::vector<std::string>& cols, const splash::ds::aligned_matrix<T>& m) {
  using h5::Id;
  h5::Lib& L = h5::lib();
  const size_t R = m.rows(), C = m.columns();
  auto name_of = [](const std::vector<std::string>& v, size_t i) { return i < v.size() ? v[i] : std::to_string(i); };
  std::vector<std::string> rn(R), cn(C);
  for (size_t i = 0; i < R; ++i) rn[i] = name_of(rows, i);
  for (size_t j = 0; j < C; ++j) cn[j] = name_of(cols, j);

  Id f(h5::check(L.H5Fcreate(path.c_str(), h5::kAccTrunc, H5P_DEFAULT, H5P_DEFAULT), "cannot create " + path),
       L.H5Fclose);
  auto str_type = [&](size_t len, H5T_cset_t cset) {
    hid_t t = L.H5Tcopy(L.C_S1);
    L.H5Tset_size(t, std::max<size_t>(len, 1));
    L.H5Tset_strpad(t, H5T_STR_NULLTERM);
    L.H5Tset_cset(t, cset);
    return t;
  };
  auto attr_str = [&](hid_t obj, const char* name, const std::string& val, H5T_cset_t cset = H5T_CSET_UTF8) {
    // PyTables sizes attribute strings without the terminating NUL and stores an empty TITLE
    // as a one-byte string with a NULL dataspace
    Id t(str_type(val.size(), cset), L.H5Tclose);
    Id s(L.H5Screate(val.empty() ? H5S_NULL : H5S_SCALAR), L.H5Sclose);
    Id a(h5::check(L.H5Acreate2(obj, name, t, s, H5P_DEFAULT, H5P_DEFAULT), std::string("attribute ") + name),
         L.H5Aclose);
    if (!val.empty()) h5::check(L.H5Awrite(a, t, val.c_str()), std::string("attribute ") + name, 0);
  };
  auto attr_i64 = [&](hid_t obj, const char* name, long long v) {
    Id s(L.H5Screate(H5S_SCALAR), L.H5Sclose);
    Id a(h5::check(L.H5Acreate2(obj, name, L.STD_I64LE, s, H5P_DEFAULT, H5P_DEFAULT), name), L.H5Aclose);
    h5::check(L.H5Awrite(a, L.NATIVE_LLONG, &v), name, 0);
  };
  auto attr_b8 = [&](hid_t obj, const char* name, unsigned char v) {
    Id s(L.H5Screate(H5S_SCALAR), L.H5Sclose);
    Id a(h5::check(L.H5Acreate2(obj, name, L.STD_B8LE, s, H5P_DEFAULT, H5P_DEFAULT), name), L.H5Aclose);
    h5::check(L.H5Awrite(a, L.NATIVE_B8, &v), name, 0);
  };
  auto array_attrs = [&](hid_t ds, const char* pd_name) {
    attr_str(ds, "CLASS", "ARRAY");
    attr_str(ds, "FLAVOR", "numpy");
    attr_str(ds, "TITLE", "");
    attr_str(ds, "VERSION", "2.4");
    if (pd_name) {
      attr_str(ds, "kind", "string");
      // pandas stores an unnamed axis as the pickled None "N." (bytes), a named one as text
      attr_str(ds, "name", pd_name, std::string(pd_name) == "N." ? H5T_CSET_ASCII : H5T_CSET_UTF8);
    }
    attr_b8(ds, "transposed", 1);
  };
  auto write_names = [&](hid_t g, const char* key, const std::vector<std::string>& v, const char* pd_name) {
    size_t w = 0;   // pandas writes name arrays one byte wider than the longest name
    for (const auto& s : v) w = std::max(w, s.size());
    ++w;
    std::vector<char> buf(std::max<size_t>(v.size(), 1) * w, 0);
    for (size_t i = 0; i < v.size(); ++i) memcpy(buf.data() + i * w, v[i].data(), v[i].size());
    Id t(str_type(w, H5T_CSET_ASCII), L.H5Tclose);
    const hsize_t n = v.size();
    Id s(L.H5Screate_simple(1, &n, nullptr), L.H5Sclose);
    Id ds(h5::check(L.H5Dcreate2(g, key, t, s, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), std::string("dataset ") + key),
          L.H5Dclose);
    if (n) h5::check(L.H5Dwrite(ds, t, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf.data()), std::string("writing ") + key, 0);
    array_attrs(ds, pd_name);
  };

  attr_str(f, "CLASS", "GROUP");
  attr_str(f, "PYTABLES_FORMAT_VERSION", "2.1");
  attr_str(f, "TITLE", "");
  attr_str(f, "VERSION", "1.0");
  const std::string gname = group.empty() ? "array" : group;
  Id g(h5::check(L.H5Gcreate2(f, gname.c_str(), H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), "group " + gname), L.H5Gclose);
  attr_str(g, "CLASS", "GROUP");
  attr_str(g, "TITLE", "");
  attr_str(g, "VERSION", "1.0");
  attr_str(g, "axis0_variety", "regular");
  attr_str(g, "axis1_variety", "regular");
  attr_str(g, "block0_items_variety", "regular");
  attr_str(g, "encoding", "UTF-8");
  attr_str(g, "errors", "strict");
  attr_i64(g, "nblocks", 1);
  attr_i64(g, "ndim", 2);
  attr_str(g, "pandas_type", "frame");
  attr_str(g, "pandas_version", "0.15.2");
  write_names(g, "axis0", cn, "N.");
  write_names(g, "axis1", rn, "genes");
  write_names(g, "block0_items", cn, "N.");
  {
    const hsize_t dims[2] = {R, C};
    Id s(L.H5Screate_simple(2, dims, nullptr), L.H5Sclose);
    const hid_t ft = std::is_same<T, float>::value ? L.IEEE_F32LE : L.IEEE_F64LE;
    Id ds(h5::check(L.H5Dcreate2(g, "block0_values", ft, s, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT),
                    "dataset block0_values"),
          L.H5Dclose);
    if (R && C) {
      std::vector<T> buf(R * C);
      for (size_t i = 0; i < R; ++i) std::copy(m.data(i), m.data(i) + C, buf.begin() + i * C);
      h5::check(L.H5Dwrite(ds, h5::native<T>(), H5S_ALL, H5S_ALL, H5P_DEFAULT, buf.data()), "writing block0_values", 0);
    }
    array_attrs(ds, nullptr);
  }
}

}  // namespace mio
