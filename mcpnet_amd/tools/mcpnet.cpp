// mcpnet: expression -> MI (B-spline or adaptive partitioning) -> diagonal -> MCP2/3/4 scores and
// the ensemble MCP.  Follows the MI, MCP and ensemble stages of apps/pipeline/mcpnet.cpp:159-700
// and the compute_* helpers of apps/pipeline/mcpnet_helper.hpp:56-638, on the GPU: MI, diagonal
// and MCP2/3/4 run as one device-resident call (mcpg_mcpnet_*), the ensemble through the
// maxmins / combo / Stouffer entry points.
// With <prefix> = <output>[_tf](_bs|_ap|_inmi) (mcpnet.cpp:265-278) the outputs are
// <prefix>.mi.<fmt>, <prefix>.mcp{2,3,4}.<fmt> and, for -m 4, <prefix>.mcp_<a>_<b>_<c>_<d>.<fmt>
// (combination ratio) and <prefix>.mcpst_<a>_<b>_<c>_<d>.<fmt> (its Stouffer combination) per
// coefficient row; <fmt> is h5 unless --format says otherwise.  With --tf-input every output
// keeps only the TF rows: the MCP products start at a TF and pass through TFs
// (compute_mcp{2,3,4}_tfs, compute_maxmins_tfs) and the ensemble's Stouffer step z-scores TF rows
// and gene columns (compute_stouffer_tf).  The evaluation stage (-g, -x, --eval: AUPR / AUROC
// against a ground truth) is not part of this build.
#include <cstdio>
#include <iomanip>
#include <set>
#include <sstream>
#include <stdexcept>
#include <type_traits>

#include "common.hpp"
#include "mcp/filter/mcp.hpp"
#include "mcp/transform/combine.hpp"
#include "mcp/transform/stouffer.hpp"

namespace {

// MI -> diagonal -> TF rows -> MCP2/3/4 in one device-resident call, in place of compute_mi /
// compute_ap_mi (mcpnet_helper.hpp:56-167), the diagonal loop (mcpnet.cpp:299-308),
// filter_local_mat_rows_by_tfs and compute_mcp{2,3,4}[_tfs] (mcpnet_helper.hpp:188-435): the
// MI and the max-min intermediates never cross the bus.
inline int mcpnet_call(mcpg_handle h, const float* x, int64_t G, int64_t S, int64_t ldx, int mim, int precision,
                       uint64_t seed, double diagonal, const float* tfs, float* mi, int64_t ldmi, float* const* o,
                       const int64_t* ldo) {
  return mcpg_mcpnet_f32(h, MCPG_HOST, x, G, S, ldx, mim, 10, 3, precision, seed, diagonal, tfs, mi, ldmi, o[0], ldo[0],
                         o[1], ldo[1], o[2], ldo[2]);
}
inline int mcpnet_call(mcpg_handle h, const double* x, int64_t G, int64_t S, int64_t ldx, int mim, int precision,
                       uint64_t seed, double diagonal, const double* tfs, double* mi, int64_t ldmi, double* const* o,
                       const int64_t* ldo) {
  return mcpg_mcpnet_f64(h, MCPG_HOST, x, G, S, ldx, mim, 10, 3, precision, seed, diagonal, tfs, mi, ldmi, o[0], ldo[0],
                         o[1], ldo[1], o[2], ldo[2]);
}

// compute_maxmins (mcpnet_helper.hpp:436-460): M1 = A(x)A, M2 = A(x)M1, M3 = M1(x)M1
inline int maxmins_call(mcpg_handle h, const float* a, int64_t la, int64_t N, float* m1, int64_t l1, float* m2, int64_t l2,
                        float* m3, int64_t l3) {
  return mcpg_maxmins_f32(h, MCPG_HOST, a, la, N, m1, l1, m2, l2, m3, l3);
}
inline int maxmins_call(mcpg_handle h, const double* a, int64_t la, int64_t N, double* m1, int64_t l1, double* m2,
                        int64_t l2, double* m3, int64_t l3) {
  return mcpg_maxmins_f64(h, MCPG_HOST, a, la, N, m1, l1, m2, l2, m3, l3);
}
template <typename T>
void compute_maxmins(const splash::ds::aligned_matrix<T>& g2g, splash::ds::aligned_matrix<T>& m1,
                     splash::ds::aligned_matrix<T>& m2, splash::ds::aligned_matrix<T>& m3) {
  const size_t N = g2g.rows();
  m1.resize(N, N);
  m2.resize(N, N);
  m3.resize(N, N);
  mcpgpu::check(maxmins_call(mcpgpu::thread_handle(), g2g.data(), g2g.ld(), N, m1.data(), m1.ld(), m2.data(), m2.ld(),
                             m3.data(), m3.ld()),
                "compute_maxmins");
}

// compute_combo (mcpnet_helper.hpp:509-571)
inline int combo_call(mcpg_handle h, const float* mi, int64_t lmi, const float* m1, int64_t l1, const float* m2,
                      int64_t l2, const float* m3, int64_t l3, int64_t M, int64_t N, const float* c, float* o,
                      int64_t lo) {
  return mcpg_combo_f32(h, MCPG_HOST, mi, lmi, m1, l1, m2, l2, m3, l3, M, N, c, o, lo);
}
inline int combo_call(mcpg_handle h, const double* mi, int64_t lmi, const double* m1, int64_t l1, const double* m2,
                      int64_t l2, const double* m3, int64_t l3, int64_t M, int64_t N, const double* c, double* o,
                      int64_t lo) {
  return mcpg_combo_f64(h, MCPG_HOST, mi, lmi, m1, l1, m2, l2, m3, l3, M, N, c, o, lo);
}
template <typename T>
void compute_combo(const splash::ds::aligned_matrix<T>& mi, const splash::ds::aligned_matrix<T>& m1,
                   const splash::ds::aligned_matrix<T>& m2, const splash::ds::aligned_matrix<T>& m3, const T* cs,
                   splash::ds::aligned_matrix<T>& out) {
  out.resize(m1.rows(), m1.columns());
  mcpgpu::check(combo_call(mcpgpu::thread_handle(), mi.data(), mi.ld(), m1.data(), m1.ld(), m2.data(), m2.ld(),
                           m3.data(), m3.ld(), m1.rows(), m1.columns(), cs, out.data(), out.ld()),
                "compute_combo");
}

// compute_stouffer (mcpnet_helper.hpp:574-584)
template <typename T>
void compute_stouffer(const splash::ds::aligned_matrix<T>& in, splash::ds::aligned_matrix<T>& out) {
  using Mat = splash::ds::aligned_matrix<T>;
  splash::pattern::ReduceTransform<Mat, splash::kernel::GaussianParamsExclude1<T, T, true>,
                                   mcp::kernel::stouffer_vector_kernel<T>, Mat>
      transformer;
  transformer(in, splash::kernel::GaussianParamsExclude1<T, T, true>(), mcp::kernel::stouffer_vector_kernel<T>(), out);
}

// compute_maxmins_tfs (mcpnet_helper.hpp:469-505), all products masked to TF intermediates:
// M1 = tg (x) gg (ttg), M2 = M1 (x) gg (tttg), M3 = M1 (x) (gg (x) gg) (ttttg)
template <typename T>
void compute_maxmins_tfs(const splash::ds::aligned_matrix<T>& t2g, const splash::ds::aligned_matrix<T>& g2g,
                         const std::vector<T>& tfs, splash::ds::aligned_matrix<T>& m1, splash::ds::aligned_matrix<T>& m2,
                         splash::ds::aligned_matrix<T>& m3) {
  using Mat = splash::ds::aligned_matrix<T>;
  using MMX = mcp::kernel::mcp2_maxmin_kernel<T, true, true>;
  const MMX mmx(tfs);
  splash::pattern::InnerProduct<Mat, MMX, Mat, false> mxgen;
  splash::pattern::InnerProduct<Mat, MMX, Mat, true> mxgen_s;
  mxgen(t2g, g2g, mmx, m1);
  mxgen(m1, g2g, mmx, m2);
  Mat m3a;
  mxgen_s(g2g, g2g, mmx, m3a);
  mxgen(m1, m3a, mmx, m3);
}

// compute_stouffer_tf (mcpnet_helper.hpp:587-638)
template <typename T>
void compute_stouffer_tf(const splash::ds::aligned_matrix<T>& in, splash::ds::aligned_matrix<T>& out) {
  out.resize(in.rows(), in.columns());
  mcpgpu::check(mcpgpu::stouffer_tf_call(mcpgpu::thread_handle(), in.data(), in.ld(), in.rows(), in.columns(),
                                         out.data(), out.ld()),
                "compute_stouffer_tf");
}

// -1 selects the DataType's own arithmetic: MCPG_FP64 for the default double run, MCPG_FP32 with
// --single (which cannot take MCPG_FP64)
template <typename T>
int precision_code(const cli::Parser& p) {
  const int c = (int)p.integer("--precision-code");
  const bool dbl = std::is_same<T, double>::value;
  if (c == -1) return dbl ? MCPG_FP64 : MCPG_FP32;
  if (c < 0 || c > MCPG_FP64) throw std::runtime_error("--precision-code must be 0, 1, 2, 3 or 4");
  if (c == MCPG_FP64 && !dbl) throw std::runtime_error("--precision-code 4 (fp64) needs the double DataType");
  return c;
}

template <typename T>
int run(const cli::Parser& p) {
  using Mat = splash::ds::aligned_matrix<T>;
  std::vector<std::string> genes, samples;
  tool::Timer t;
  Mat input = tool::load_input<T>(p, genes, samples);
  printf("Load data in %f sec\n", t.seconds());
  // transcription factors (mcpnet.cpp:201-213): the outputs keep only the TF rows
  const std::string tf_input = p.str("--tf-input");
  const bool tfonly = !tf_input.empty();
  std::vector<T> tfs;
  std::vector<std::string> tf_names;
  if (tfonly) {
    tool::load_tfs<T, double>(tf_input, genes, tfs, tf_names);
    if (tf_names.empty()) {
      printf("ERROR: no transcription factors in the specified file");
      return 0;
    }
  }
  const std::vector<std::string>& row_names = tfonly ? tf_names : genes;
  const bool mi_input = p.seen("--mi-input");
  const int mim = (int)p.integer("--mi-method");
  if (mim < 0 || mim > 1) throw std::runtime_error("--mi-method must be 0 (B-spline) or 1 (adaptive partitioning)");
  // output names: <output>[_tf](_bs|_ap|_inmi).<stage>.h5 (mcpnet.cpp:265-278)
  const std::string prefix =
      p.str("--output") + (tfonly ? "_tf" : "") + (mi_input ? "_inmi" : mim == 0 ? "_bs" : "_ap");
  const std::string fmt = "." + p.str("--format");
  const size_t G = input.rows();
  if (mi_input && input.columns() != G) throw std::runtime_error("an MI input must be square");
  if (tfonly && G != genes.size()) throw std::runtime_error("the input must have one row per gene");
  std::set<int> comps;
  for (const auto& m : p.all("--method")) comps.insert(std::stoi(m));
  if (comps.empty()) comps.insert(2);
  for (int m : comps)
    if (m < 1 || m > 4) throw std::runtime_error("-m accepts 1 (MCP2), 2 (MCP3), 3 (MCP4), 4 (ensemble MCP)");

  t = tool::Timer();
  Mat mi(G, G);
  Mat dratio[3];
  T* outs[3] = {nullptr, nullptr, nullptr};
  int64_t ldo[3] = {0, 0, 0};
  for (int m = 1; m <= 3; ++m)
    if (comps.count(m)) {
      dratio[m - 1].resize(tfonly ? tf_names.size() : G, G);
      outs[m - 1] = dratio[m - 1].data();
      ldo[m - 1] = dratio[m - 1].ld();
    }
  mcpgpu::check(mcpnet_call(mcpgpu::thread_handle(), input.data(), G, mi_input ? 0 : input.columns(), input.ld(),
                            mi_input ? -1 : mim, precision_code<T>(p), (uint64_t)p.integer("--noise-seed"),
                            p.real("--diagonal"), tfonly ? tfs.data() : nullptr, mi.data(), mi.ld(), outs, ldo),
                "mcpnet pipeline");
  printf("Computed,MI and MCP,,%f,sec\n", t.seconds());
  Mat mi_tf;
  if (tfonly) mi_tf = tool::filter_rows_by_tfs(mi, tfs);   // mcpnet.cpp:316-324
  const Mat& mi_rows = tfonly ? mi_tf : mi;
  t = tool::Timer();
  if (!mi_input) mio::write_matrix(prefix + ".mi" + fmt, row_names, genes, mi_rows);   // mcpnet.cpp:353-356
  for (int m = 1; m <= 3; ++m)
    if (comps.count(m)) mio::write_matrix(prefix + ".mcp" + std::to_string(m + 1) + fmt, row_names, genes, dratio[m - 1]);
  printf("Wrote,MI and MCP,,%f,sec\n", t.seconds());
  if (comps.count(4)) {   // ensemble MCP (mcpnet.cpp:518-700, the branch without evaluation)
    t = tool::Timer();
    Mat coeffs;
    if (!p.str("--coefficients").empty()) {
      std::vector<std::string> rdummy, cdummy;
      coeffs = mio::read_matrix<T>(p.str("--coefficients"), rdummy, cdummy);
      if (coeffs.columns() < 4) throw std::runtime_error("the coefficient file needs 4 columns");
    } else {
      coeffs.resize(1, 4);
      for (size_t j = 0; j < 4; ++j) coeffs.at(0, j) = static_cast<T>(0.25);
    }
    printf("Loaded,%zux%zu coefficients,,%f,sec\n", coeffs.rows(), coeffs.columns(), t.seconds());
    t = tool::Timer();
    Mat dmaxmin1, dmaxmin2, dmaxmin3;
    if (tfonly) compute_maxmins_tfs(mi_tf, mi, tfs, dmaxmin1, dmaxmin2, dmaxmin3);
    else compute_maxmins(mi, dmaxmin1, dmaxmin2, dmaxmin3);
    printf("Computed,MAXMIN 1 2 3,,%f,sec\n", t.seconds());
    for (size_t i = 0; i < coeffs.rows(); ++i) {
      const T* c = coeffs.data(i);
      if (c[0] == 1 && c[1] == 0 && c[2] == 0 && c[3] == 0) continue;
      std::ostringstream tag;
      tag << std::fixed << std::setprecision(3) << c[0] << "_" << c[1] << "_" << c[2] << "_" << c[3];
      t = tool::Timer();
      Mat dratio, dratiost;
      compute_combo(mi_rows, dmaxmin1, dmaxmin2, dmaxmin3, c, dratio);
      printf("Computed,Combo,(%s),%f,sec\n", tag.str().c_str(), t.seconds());
      mio::write_matrix(prefix + ".mcp_" + tag.str() + fmt, row_names, genes, dratio);
      t = tool::Timer();
      if (tfonly) compute_stouffer_tf(dratio, dratiost);
      else compute_stouffer(dratio, dratiost);
      printf("Computed,Combo Stouffer,(%s),%f,sec\n", tag.str().c_str(), t.seconds());
      mio::write_matrix(prefix + ".mcpst_" + tag.str() + fmt, row_names, genes, dratiost);
    }
  }
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  cli::Parser p("mcpnet");
  tool::common_options(p);
  p.option("--mi-method", "MI Algorithm: 0 = Bspline, 1 (default) = Adaptive Partitioning", true, "1");
  p.option("--mi-input", "input file holds a precomputed MI matrix", false);
  p.option("-m,--method", "Algorithm: MCP2=1, MCP3=2, MCP4=3, EnsembleMCP=4 (one or more)", true, "");
  p.option("-f,--coefficients", "file with combo coefficients (rows of 4).  For ensemble MCP only.", true, "");
  p.option("--clamped", "output is clamped (no effect, as in the reference)", false);
  p.option("--diagonal", "MI diagonal value; negative keeps the MI diagonal", true, "0");
  p.option("--format", "output format: h5 (as the reference), exp, csv or tsv", true, "h5");
  p.option("--precision-code",
           "B-spline contraction: 0 fp32, 1 bf16, 2 bf16x3, 3 i8x3, 4 fp64 (default: 4, 0 with --single)", true, "-1");
  p.option("--noise-seed", "seed of the U[0,1e-5) rank jitter", true, "1296388676612");
  p.option("--tf-input", "transcription factor file, 1 per line.", true, "");
  try {
    p.parse(argc, argv);
    if (p.str("--output").empty()) throw std::runtime_error("no output prefix: give -o PREFIX");
    tool::Timer t;
    int rc = p.seen("--single") ? run<float>(p) : run<double>(p);
    printf("Total Computation in %f sec\n", t.seconds());
    return rc;
  } catch (const std::exception& e) {
    fprintf(stderr, "mcpnet: %s\n", e.what());
    return 1;
  }
}
