// mcpnet: expression -> MI (B-spline or adaptive partitioning) -> diagonal -> MCP2/3/4 scores.
// Follows the MI and MCP stages of apps/pipeline/mcpnet.cpp:159-515 and the compute_* helpers of
// apps/pipeline/mcpnet_helper.hpp:56-364, on the GPU.  Outputs are <prefix>.mi.<fmt> and
// <prefix>.mcp{2,3,4}.<fmt>.  The ensemble (-f coefficients), Stouffer and evaluation (-g, -x,
// --eval) stages of the reference are not part of this build.
#include <cstdio>
#include <set>

#include "common.hpp"
#include "mcp/correlation/AdaptivePartitioningMI.hpp"
#include "mcp/correlation/BSplineMI.hpp"
#include "mcp/filter/mcp.hpp"
#include "mcp/transform/combine.hpp"
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

template <typename T>
int run(const cli::Parser& p) {
  using Mat = splash::ds::aligned_matrix<T>;
  std::vector<std::string> genes, samples;
  tool::Timer t;
  Mat input = tool::load_input<T>(p, genes, samples);
  printf("Load data in %f sec\n", t.seconds());
  const std::string prefix = p.str("--output"), fmt = "." + p.str("--format");
  Mat mi;
  if (p.seen("--mi-input")) {
    mi = input;
  } else {
    t = tool::Timer();
    const int mim = (int)p.integer("--mi-method");
    if (mim == 0) compute_mi(input, 10, (int)p.integer("--precision-code"), mi);
    else compute_ap_mi(input, (uint64_t)p.integer("--noise-seed"), mi);
    printf("Computed,MI,,%f,sec\n", t.seconds());
    mio::write_matrix(prefix + ".mi" + fmt, genes, genes, mi);
  }
  const double diagonal = p.real("--diagonal");
  if (diagonal >= 0.0)   // mcpnet.cpp:299-308
    for (size_t i = 0; i < std::min(mi.rows(), mi.columns()); ++i) mi.at(i, i) = static_cast<T>(diagonal);
  std::set<int> comps;
  for (const auto& m : p.all("--method")) comps.insert(std::stoi(m));
  if (comps.empty()) comps.insert(2);
  for (int m : comps) {
    if (m < 1 || m > 3) throw std::runtime_error("-m accepts 1 (MCP2), 2 (MCP3), 3 (MCP4)");
    t = tool::Timer();
    Mat dratio;
    compute_mcp(mi, m + 1, dratio);
    printf("Computed,MCP%d,,%f,sec\n", m + 1, t.seconds());
    mio::write_matrix(prefix + ".mcp" + std::to_string(m + 1) + fmt, genes, genes, dratio);
  }
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  cli::Parser p("mcpnet");
  tool::common_options(p);
  p.option("--mi-method", "MI Algorithm: 0 = Bspline, 1 (default) = Adaptive Partitioning", true, "1");
  p.option("--mi-input", "input file holds a precomputed MI matrix", false);
  p.option("-m,--method", "MCP2=1, MCP3=2, MCP4=3 (one or more)", true, "");
  p.option("--clamped", "output is clamped (no effect, as in the reference)", false);
  p.option("--diagonal", "MI diagonal value; negative keeps the MI diagonal", true, "0");
  p.option("--format", "output format: exp, csv or tsv", true, "exp");
  p.option("--precision-code", "B-spline contraction: 0 fp32, 1 bf16, 2 bf16x3", true, "0");
  p.option("--noise-seed", "seed of the U[0,1e-5) rank jitter", true, "1296388676612");
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
