// convert: reads a matrix (EXP, CSV/TSV, HDF5 or a random one) and writes it in the format the
// output path names.  Follows apps/transform/convert.cpp:34-84 of the reference: no transpose, the
// HDF5 output goes to group "array" (convert.cpp:72).  Host code only; it needs no GPU.
#include <cstdio>

#include "common.hpp"

namespace {

template <typename DataType>
int run(const cli::Parser& p) {
  std::vector<std::string> genes, samples;
  tool::Timer t;
  splash::ds::aligned_matrix<DataType> input = tool::load_input<DataType>(p, genes, samples);
  printf("Load data from %s in %f sec\n", p.str("--input").c_str(), t.seconds());
  printf("No. GENES    : %zu\nNo. SAMPLES  : %zu\n", input.rows(), input.columns());
  t = tool::Timer();
  mio::write_matrix(p.str("--output"), genes, samples, input);
  printf("converted to %s in %f sec\n", p.str("--output").c_str(), t.seconds());
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  cli::Parser p("convert");
  tool::common_options(p);
  try {
    p.parse(argc, argv);
    if (p.str("--output").empty()) throw std::runtime_error("no output: give -o FILE");
    return p.seen("--single") ? run<float>(p) : run<double>(p);
  } catch (const std::exception& e) {
    fprintf(stderr, "convert: %s\n", e.what());
    return 1;
  }
}
