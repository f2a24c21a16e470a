// dpi: data-processing-inequality filter of an MI matrix.  Follows apps/filter/dpi.cpp:96-320 of
// the reference: optional diagonal, optional transcription-factor list (dpi_tf_kernel, then only
// the TF rows are written), the DPI mask over all gene pairs and its application (removed edges
// become 0), with the mask and its application running on the GPU (mcpg_dpi_*).
#include <algorithm>
#include <cstdio>
#include <fstream>
#include <limits>
#include <unordered_set>

#include "common.hpp"
#include "mcp/filter/dpi.hpp"

namespace {

template <typename DataType>
int run(const cli::Parser& p) {
  using MatrixType = splash::ds::aligned_matrix<DataType>;
  std::vector<std::string> genes, samples;
  tool::Timer t;
  MatrixType input = tool::load_input<DataType>(p, genes, samples);
  // dpi.cpp:123-125: transposed when the shape matches (samples x genes), square inputs included
  if (input.rows() == samples.size() && input.columns() == genes.size()) input = input.local_transpose();
  printf("Load data in %f sec\n", t.seconds());
  const double diagonal = p.real("--diagonal");
  if (diagonal >= 0.0)   // dpi.cpp:130-135
    for (size_t i = 0; i < std::min(input.rows(), input.columns()); ++i) input.at(i, i) = static_cast<DataType>(diagonal);
  if (input.rows() != input.columns()) throw std::runtime_error("input matrix should be square");
  const double tol_in = p.real("--tolerance");
  if (!(tol_in >= 0.0 && tol_in <= 1.0)) throw std::runtime_error("--tolerance must be in [0, 1]");
  const DataType tolerance = static_cast<DataType>(tol_in);

  // transcription factors (dpi.cpp:165-215): max for a TF, lowest otherwise
  std::vector<DataType> tfs;
  std::vector<std::string> tf_names;
  const std::string tf_input = p.str("--tf-input");
  if (!tf_input.empty()) {
    tool::load_tfs(tf_input, genes, tfs, tf_names);
    if (tf_names.empty()) return 0;
  }

  t = tool::Timer();
  const size_t N = input.rows();
  MatrixType dout(N, N);
  int64_t removed = 0;
  mcpgpu::check(mcpgpu::dpi_call(mcpgpu::thread_handle(), input.data(), input.ld(), N, tolerance,
                                 tfs.empty() ? nullptr : tfs.data(), dout.data(), dout.ld(), nullptr, 0, &removed),
                "dpi");
  printf("Generate and apply mask (%zu x %zu) in %f sec\n", N, N, t.seconds());
  printf("Mask removing %lld elements \n", (long long)removed);

  MatrixType output;
  if (!tfs.empty()) {   // keep the TF rows (dpi.cpp:288-302)
    output.resize(tf_names.size(), N);
    for (size_t i = 0, j = 0; i < N; ++i)
      if (tfs[i] > 0) {
        std::copy(dout.data(i), dout.data(i) + N, output.data(j));
        ++j;
      }
  } else {
    output = dout;
  }
  printf("output size: %zu x %zu\n", output.rows(), output.columns());
  mio::write_matrix(p.str("--output"), tfs.empty() ? genes : tf_names, samples, output);
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  cli::Parser p("dpi");
  tool::common_options(p);
  p.option("-l,--tolerance", "value comparison tolerance.", true, "0.1");
  p.option("-f,--tf-input", "transcription factor file, 1 per line.", true, "");
  p.option("--tf-gene-transition", "transition from tf-tf to gene-gene (accepted; as in the reference it changes nothing)",
           true, "-1");
  p.option("--diagonal", "Input MI matrix diagonal should be set as this value. default 0. if negative, use original MI",
           true, "0");
  try {
    p.parse(argc, argv);
    if (p.str("--output").empty()) throw std::runtime_error("no output: give -o FILE");
    tool::Timer t;
    int rc = p.seen("--single") ? run<float>(p) : run<double>(p);
    printf("Total Computation in %f sec\n", t.seconds());
    return rc;
  } catch (const std::exception& e) {
    fprintf(stderr, "dpi: %s\n", e.what());
    return 1;
  }
}
