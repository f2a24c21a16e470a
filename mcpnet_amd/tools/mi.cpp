// mi: all-pairs mutual information, B-spline (-m 0) or adaptive partitioning (-m 1, default).
// Follows apps/correlation/mi.cpp:69-265 of the reference, with every pattern call running on the
// GPU.  Unlike the reference (whose B-spline InnerProduct call is commented out at mi.cpp:251),
// -m 0 computes the MI.
#include <cstdio>
#include <iostream>
#include <stdexcept>
#include <type_traits>

#include "common.hpp"
#include "mcp/correlation/AdaptivePartitioningMI.hpp"
#include "mcp/correlation/BSplineMI.hpp"
#include "splash/transform/rank.hpp"

namespace {

// "" selects the DataType's own arithmetic: fp64 for the default double run, fp32 with --single
template <typename DataType>
int precision_of(const std::string& s) {
  const bool dbl = std::is_same<DataType, double>::value;
  if (s.empty()) return dbl ? MCPG_FP64 : MCPG_FP32;
  if (s == "fp32") return MCPG_FP32;
  if (s == "bf16") return MCPG_BF16;
  if (s == "bf16x3") return MCPG_BF16X3;
  if (s == "i8x3") return MCPG_I8X3;
  if (s == "fp64") {
    if (!dbl) throw std::runtime_error("--precision fp64 needs the double DataType (no --single)");
    return MCPG_FP64;
  }
  throw std::runtime_error("--precision must be fp32, bf16, bf16x3, i8x3 or fp64");
}

template <typename DataType>
int run(const cli::Parser& p) {
  using Mat = splash::ds::aligned_matrix<DataType>;
  std::vector<std::string> genes, samples;
  tool::Timer t;
  Mat input = tool::load_input<DataType>(p, genes, samples);
  // mi.cpp:95-96: a matrix whose shape matches (samples x genes) is transposed; a square input,
  // whose gene and sample counts agree, is transposed as well
  if (input.rows() == samples.size() && input.columns() == genes.size()) input = input.local_transpose();
  printf("Load data in %f sec\n", t.seconds());
  printf("No. GENES    : %zu\nNo. SAMPLES  : %zu\n", input.rows(), input.columns());
  const int method = (int)p.integer("--method");
  Mat output;
  if (method == 1) {
    t = tool::Timer();
    Mat noisy(input.rows(), input.columns());
    using NoiseKernelType = mcp::kernel::add_noise<DataType, std::uniform_real_distribution<DataType>,
                                                   std::default_random_engine>;
    splash::kernel::random_number_generator<> gen(p.integer("--noise-seed"));
    NoiseKernelType noise_adder(gen, 0.0, 0.00001);
    noise_adder(input, noisy);
    printf("Noise added in %f sec\n", t.seconds());
    t = tool::Timer();
    using RankKernelType = splash::kernel::Rank<DataType, size_t, 0, false>;
    splash::pattern::Transform<Mat, RankKernelType, splash::ds::aligned_matrix<size_t>> normalizer;
    splash::ds::aligned_matrix<size_t> ranked(noisy.rows(), noisy.columns());
    normalizer(noisy, RankKernelType(), ranked);
    printf("Rank Transformed in %f sec\n", t.seconds());
    t = tool::Timer();
    using MIKernelType = mcp::kernel::AdaptivePartitionRankMIKernel2<size_t, DataType, 0>;
    splash::pattern::InnerProduct<splash::ds::aligned_matrix<size_t>, MIKernelType, Mat, true> adaptive_mi;
    output.resize(ranked.rows(), ranked.rows());
    adaptive_mi(ranked, ranked, MIKernelType(), output);
    printf("Compute Adpative MI in %f sec\n", t.seconds());
  } else if (method == 0) {
    t = tool::Timer();
    splash::pattern::GlobalTransform<Mat, mcp::kernel::MinMaxScale<DataType>, Mat> scaler;
    Mat scaled;
    scaler(input, mcp::kernel::MinMaxScale<DataType>(), scaled);
    printf("Scaled Partitioned in %f sec\n", t.seconds());
    t = tool::Timer();
    const int bins = (int)p.integer("--bins"), order = (int)p.integer("--spline_order");
    const int S = (int)input.columns();
    splash::pattern::Transform<Mat, mcp::kernel::BSplineWeightsKernel<DataType>, Mat> bspline_weights;
    Mat weighted(scaled.rows(), (size_t)S * bins + 1);
    bspline_weights(scaled, mcp::kernel::BSplineWeightsKernel<DataType>(bins, order, S), weighted);
    printf("Weight Partitioned in %f sec\n", t.seconds());
    t = tool::Timer();
    splash::pattern::InnerProduct<Mat, mcp::kernel::BSplineMIKernel<DataType>, Mat, true> mi_proc;
    output.resize(input.rows(), input.rows());
    mi_proc(weighted, weighted,
            mcp::kernel::BSplineMIKernel<DataType>(bins, S, precision_of<DataType>(p.str("--precision"))),
            output);
    printf("Compute MI in %f sec\n", t.seconds());
  } else {
    throw std::runtime_error("--method must be 0 (B-spline) or 1 (adaptive partitioning)");
  }
  t = tool::Timer();
  mio::write_matrix(p.str("--output"), genes, genes, output);
  printf("Output in %f sec\n", t.seconds());
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  cli::Parser p("mi");
  tool::common_options(p);
  p.option("-b,--bins", "No. of bins", true, "10");
  p.option("-l,--spline_order", "Spline Order", true, "3");
  p.option("-m,--method", "Algorithm: B-spline MI=0, Adaptive partitioning=1", true, "1");
  p.option("--precision",
           "B-spline contraction arithmetic: fp32, bf16, bf16x3, i8x3 or fp64 (default: fp64, fp32 with --single)",
           true, "");
  p.option("--noise-seed", "seed of the U[0,1e-5) rank jitter", true, "1296388676612");
  try {
    p.parse(argc, argv);
    if (p.str("--output").empty()) throw std::runtime_error("no output: give -o FILE");
    tool::Timer t;
    int rc = p.seen("--single") ? run<float>(p) : run<double>(p);
    printf("Total Computation in %f sec\n", t.seconds());
    return rc;
  } catch (const std::exception& e) {
    fprintf(stderr, "mi: %s\n", e.what());
    return 1;
  }
}
