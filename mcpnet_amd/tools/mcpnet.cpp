// mcpnet: expression -> MI (B-spline or adaptive partitioning) -> diagonal -> MCP2/3/4 scores and
// the ensemble MCP.  Follows the MI, MCP and ensemble stages of apps/pipeline/mcpnet.cpp:159-700
// and the compute_* helpers of apps/pipeline/mcpnet_helper.hpp:56-584, on the GPU.  Outputs are
// <prefix>.mi.<fmt>, <prefix>.mcp{2,3,4}.<fmt> and, for -m 4, <prefix>.mcp_<a>_<b>_<c>_<d>.<fmt>
// (combination ratio) and <prefix>.mcpst_<a>_<b>_<c>_<d>.<fmt> (its Stouffer combination) per
// coefficient row.  The evaluation stage (-g, -x, --eval: AUPR / AUROC against a ground truth)
// is not part of this build.
#include <cstdio>
#include <iomanip>
#include <set>
#include <sstream>

#include "common.hpp"
#include "mcp/correlation/AdaptivePartitioningMI.hpp"
#include "mcp/correlation/BSplineMI.hpp"
#include "mcp/filter/mcp.hpp"
#include "mcp/transform/combine.hpp"
#include "mcp/transform/stouffer.hpp"
#include "splash/kernel/rank.hpp"

namespace {

// compute_mi (mcpnet_helper.hpp:56-99)
template <typename T>
void compute_mi(const splash::ds::aligned_matrix<T>& input, int bins, int precision, splash::ds::aligned_matrix<T>& dmi) {
  using Mat = splash::ds::aligned_matrix<T>;
  splash::pattern::GlobalTransform<Mat, mcp::kernel::MinMaxScale<T>, Mat> scaler;
  Mat dscaled(input.rows(), input.columns());
  scaler(input, mcp::kernel::MinMaxScale<T>(), dscaled);
  const int S = (int)input.columns();
  splash::pattern::Transform<Mat, mcp::kernel::BSplineWeightsKernel<T>, Mat> bspline_weights;
  Mat weighted(dscaled.rows(), (size_t)S * bins + 1);
  bspline_weights(dscaled, mcp::kernel::BSplineWeightsKernel<T>(bins, 3, S), weighted);
  splash::pattern::InnerProduct<Mat, mcp::kernel::BSplineMIKernel<T>, Mat, true> bspline_mi;
  bspline_mi(weighted, weighted, mcp::kernel::BSplineMIKernel<T>(bins, S, precision), dmi);
}

// compute_ap_mi (mcpnet_helper.hpp:102-167), with noise
template <typename T>
void compute_ap_mi(const splash::ds::aligned_matrix<T>& input, uint64_t seed, splash::ds::aligned_matrix<T>& dmi) {
  using Mat = splash::ds::aligned_matrix<T>;
  Mat noisy(input.rows(), input.columns());
  mcp::kernel::add_noise<T> noise_adder(seed);
  noise_adder(input, noisy);
  using RankKernelType = splash::kernel::Rank<T, size_t, 0, false>;
  splash::pattern::Transform<Mat, RankKernelType, splash::ds::aligned_matrix<size_t>> normalizer;
  splash::ds::aligned_matrix<size_t> ranked(noisy.rows(), noisy.columns());
  normalizer(noisy, RankKernelType(), ranked);
  using MIKernelType = mcp::kernel::AdaptivePartitionRankMIKernel2<size_t, T, 0>;
  splash::pattern::InnerProduct<splash::ds::aligned_matrix<size_t>, MIKernelType, Mat, true> adaptive_mi;
  adaptive_mi(ranked, ranked, MIKernelType(), dmi);
}

// compute_mcp{2,3,4} (mcpnet_helper.hpp:188-205, :247-277, :335-364)
template <typename T>
void compute_mcp(const splash::ds::aligned_matrix<T>& g2g, int k, splash::ds::aligned_matrix<T>& dmcp) {
  using Mat = splash::ds::aligned_matrix<T>;
  using MXKernelType = mcp::kernel::mcp2_maxmin_kernel<T, false, true>;
  splash::pattern::InnerProduct<Mat, MXKernelType, Mat, true> mxgen_s;
  splash::pattern::InnerProduct<Mat, MXKernelType, Mat, false> mxgen;
  Mat maxmin1, dmaxmin;
  mxgen_s(g2g, g2g, MXKernelType(), maxmin1);
  if (k == 2) dmaxmin = maxmin1;
  else if (k == 3) mxgen(g2g, maxmin1, MXKernelType(), dmaxmin);
  else mxgen_s(maxmin1, maxmin1, MXKernelType(), dmaxmin);
  splash::pattern::BinaryOp<Mat, Mat, mcp::kernel::ratio_kernel<T, T>, Mat> gen;
  const auto part = splash::utils::partition<size_t>::make_partition(dmaxmin.rows());
  gen(g2g, part, dmaxmin, part, mcp::kernel::ratio_kernel<T, T>(), dmcp, part);
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

// compute_maxmin{1,2,3}_tfs (mcpnet_helper.hpp:208-232, :280-321, :366-424): the max-min product
// of a path of k legs starting at a TF (rows of t2g), with tf_gene_transition choosing the legs
// whose intermediate genes must be TFs (the masked kernel).
template <typename T>
void compute_maxmin_tfs(const splash::ds::aligned_matrix<T>& t2g, const splash::ds::aligned_matrix<T>& g2g, int k,
                        int transition, const std::vector<T>& tfs, splash::ds::aligned_matrix<T>& dmaxmin) {
  using Mat = splash::ds::aligned_matrix<T>;
  using MX = mcp::kernel::mcp2_maxmin_kernel<T, false, true>;
  using MMX = mcp::kernel::mcp2_maxmin_kernel<T, true, true>;
  if (transition < 0 || transition > k - 1) throw std::runtime_error("tf_gene_transition out of range");
  const MX mx;
  const MMX mmx(tfs);
  splash::pattern::InnerProduct<Mat, MX, Mat, true> mxgen_s;
  splash::pattern::InnerProduct<Mat, MX, Mat, false> mxgen;
  splash::pattern::InnerProduct<Mat, MMX, Mat, true> mmxgen_s;
  splash::pattern::InnerProduct<Mat, MMX, Mat, false> mmxgen;
  if (k == 2) {
    if (transition == 1) mmxgen(t2g, g2g, mmx, dmaxmin);   // tfs-tfs-genes
    else mxgen(t2g, g2g, mx, dmaxmin);                     // tfs-genes-genes
  } else if (k == 3) {
    Mat maxmin1;
    if (transition == 2) mmxgen_s(g2g, g2g, mmx, maxmin1);
    else mxgen_s(g2g, g2g, mx, maxmin1);
    if (transition == 2 || transition == 1) mmxgen(t2g, maxmin1, mmx, dmaxmin);
    else mxgen(t2g, maxmin1, mx, dmaxmin);
  } else {
    Mat maxmin1, maxmin2;
    if (transition == 3) mmxgen_s(g2g, g2g, mmx, maxmin1);
    else mxgen_s(g2g, g2g, mx, maxmin1);
    if (transition == 0) mxgen(t2g, g2g, mx, maxmin2);
    else mmxgen(t2g, g2g, mmx, maxmin2);
    if (transition == 2 || transition == 3) mmxgen(maxmin2, maxmin1, mmx, dmaxmin);
    else mxgen(maxmin2, maxmin1, mx, dmaxmin);
  }
}

// compute_mcp{2,3,4}_tfs (mcpnet_helper.hpp:235-244, :325-333, :427-435), called by mcpnet with
// tf_gene_transition = k - 1 (mcpnet.cpp:372-376, :421-425, :468-472)
template <typename T>
void compute_mcp_tfs(const splash::ds::aligned_matrix<T>& t2g, const splash::ds::aligned_matrix<T>& g2g, int k,
                     const std::vector<T>& tfs, splash::ds::aligned_matrix<T>& dmcp) {
  using Mat = splash::ds::aligned_matrix<T>;
  Mat dmaxmin;
  compute_maxmin_tfs(t2g, g2g, k, k - 1, tfs, dmaxmin);
  splash::pattern::BinaryOp<Mat, Mat, mcp::kernel::ratio_kernel<T, T>, Mat> gen;
  const auto part = splash::utils::partition<size_t>::make_partition(dmaxmin.rows());
  gen(t2g, part, dmaxmin, part, mcp::kernel::ratio_kernel<T, T>(), dmcp, part);
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
  // output names: <output>[_tf](_bs|_ap|_inmi).<stage>.h5 (mcpnet.cpp:265-278)
  const std::string prefix =
      p.str("--output") + (tfonly ? "_tf" : "") + (mi_input ? "_inmi" : mim == 0 ? "_bs" : "_ap");
  const std::string fmt = "." + p.str("--format");
  Mat mi;
  if (mi_input) {
    mi = input;
  } else {
    t = tool::Timer();
    if (mim == 0) compute_mi(input, 10, (int)p.integer("--precision-code"), mi);
    else compute_ap_mi(input, (uint64_t)p.integer("--noise-seed"), mi);
    printf("Computed,MI,,%f,sec\n", t.seconds());
  }
  const double diagonal = p.real("--diagonal");
  if (diagonal >= 0.0)   // mcpnet.cpp:299-308
    for (size_t i = 0; i < std::min(mi.rows(), mi.columns()); ++i) mi.at(i, i) = static_cast<T>(diagonal);
  if (tfonly && mi.rows() != genes.size()) throw std::runtime_error("the MI matrix must have one row per gene");
  Mat mi_tf;
  if (tfonly) mi_tf = tool::filter_rows_by_tfs(mi, tfs);   // mcpnet.cpp:316-324
  const Mat& mi_rows = tfonly ? mi_tf : mi;
  if (!mi_input) {   // written after the diagonal is set (mcpnet.cpp:353-356)
    t = tool::Timer();
    mio::write_matrix(prefix + ".mi" + fmt, row_names, genes, mi_rows);
    printf("Wrote,MI,,%f,sec\n", t.seconds());
  }
  std::set<int> comps;
  for (const auto& m : p.all("--method")) comps.insert(std::stoi(m));
  if (comps.empty()) comps.insert(2);
  for (int m : comps) {
    if (m < 1 || m > 4) throw std::runtime_error("-m accepts 1 (MCP2), 2 (MCP3), 3 (MCP4), 4 (ensemble MCP)");
    if (m == 4) continue;
    t = tool::Timer();
    Mat dratio;
    if (tfonly) compute_mcp_tfs(mi_tf, mi, m + 1, tfs, dratio);
    else compute_mcp(mi, m + 1, dratio);
    printf("Computed,MCP%d,,%f,sec\n", m + 1, t.seconds());
    mio::write_matrix(prefix + ".mcp" + std::to_string(m + 1) + fmt, row_names, genes, dratio);
  }
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
  p.option("--precision-code", "B-spline contraction: 0 fp32, 1 bf16, 2 bf16x3", true, "0");
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
