// mcp: MCP2 (-m 1), MCP3 (-m 2), MCP4 (-m 3) scores of an MI matrix, or the MCP score from two
// partial max-min matrices (-m 5).  Follows apps/filter/mcp.cpp:139-672 of the reference,
// including the transcription-factor variants (-f, --tf-gene-transition), with every max-min,
// ratio and max pattern running on the GPU.  Differences from the reference, on purpose:
//   * -m 5 reads --second for the second partial (the reference re-reads --first, mcp.cpp:192);
//   * --clamped is accepted and, as in the reference (mcp.cpp:391-392), changes nothing.
#include <algorithm>
#include <cstdio>
#include <fstream>
#include <limits>
#include <unordered_set>

#include "common.hpp"
#include "mcp/filter/mcp.hpp"
#include "mcp/transform/combine.hpp"

namespace {

enum Method { MCP2 = 1, MCP3 = 2, MCP4 = 3, BIN_MCP = 5 };

template <typename DataType>
int run(const cli::Parser& p) {
  using MatrixType = splash::ds::aligned_matrix<DataType>;
  using MXKernelType = mcp::kernel::mcp2_maxmin_kernel<DataType, false, true>;
  using MaskedMXKernelType = mcp::kernel::mcp2_maxmin_kernel<DataType, true, true>;
  const int compute = (int)p.integer("--method");
  const std::string tf_input = p.str("--tf-input");
  const int transition = (int)p.integer("--tf-gene-transition");
  if (transition < -1 || transition > 3) throw std::runtime_error("--tf-gene-transition must be in [-1, 3]");
  std::vector<std::string> mm_out = p.all("--maxmin-output");

  std::vector<std::string> genes, samples, g1, s1, g2, s2;
  tool::Timer t;
  MatrixType input = tool::load_input<DataType>(p, genes, samples);
  MatrixType input1, input2;
  if (compute == BIN_MCP) {
    input1 = mio::read_matrix<DataType>(p.str("--first"), g1, s1, tool::h5_keys(p));
    input2 = mio::read_matrix<DataType>(p.str("--second"), g2, s2, tool::h5_keys(p));
  }
  printf("Load data in %f sec\n", t.seconds());

  // --diagonal: set the diagonal of every input (mcp.cpp:202-221)
  std::vector<std::string> diags = p.all("--diagonal");
  if (!diags.empty()) {
    const DataType d = static_cast<DataType>(std::stod(diags[0]));
    for (MatrixType* m : {&input, &input1, &input2})
      for (size_t i = 0; i < std::min(m->rows(), m->columns()); ++i) m->at(i, i) = d;
  }
  if (input.rows() != input.columns()) throw std::runtime_error("input matrix should be square");

  // transcription factors (mcp.cpp:235-315): mask = max for TF genes, lowest otherwise
  std::vector<DataType> tfs;
  std::vector<std::string> tf_names;
  if (!tf_input.empty()) {
    tool::load_tfs(tf_input, genes, tfs, tf_names);
    if (tf_names.empty()) return 0;
  }
  const bool tfonly = !tf_input.empty();
  MatrixType tfs_to_genes(1, input.columns());
  MatrixType& genes_to_genes = input;
  if (tfonly) tfs_to_genes = tool::filter_rows_by_tfs(input, tfs);

  MXKernelType mxkernel;
  MaskedMXKernelType maskedmxkernel(tfs);
  splash::pattern::InnerProduct<MatrixType, MXKernelType, MatrixType, true> mxgen_s;
  splash::pattern::InnerProduct<MatrixType, MXKernelType, MatrixType, false> mxgen;
  splash::pattern::InnerProduct<MatrixType, MaskedMXKernelType, MatrixType, true> mmxgen_s;
  splash::pattern::InnerProduct<MatrixType, MaskedMXKernelType, MatrixType, false> mmxgen;
  splash::pattern::GlobalBinaryOp<MatrixType, MatrixType, mcp::kernel::ratio_kernel<DataType, DataType>, MatrixType> tolgen;
  const mcp::kernel::ratio_kernel<DataType, DataType> ratio;
  MatrixType output, maxmin_final;
  const MatrixType& rows_in = tfonly ? tfs_to_genes : genes_to_genes;

  t = tool::Timer();
  if (compute == MCP2) {
    MatrixType dmaxmin1;
    if (tfonly) {
      if (transition == 1) mmxgen(tfs_to_genes, genes_to_genes, maskedmxkernel, dmaxmin1);   // tfs-tfs-genes
      else mxgen(tfs_to_genes, genes_to_genes, mxkernel, dmaxmin1);                           // tfs-genes-genes
    } else {
      mxgen_s(genes_to_genes, genes_to_genes, mxkernel, dmaxmin1);
    }
    maxmin_final = dmaxmin1;
  } else if (compute == MCP3) {
    MatrixType maxmin1, maxmin2;
    if (tfonly && transition == 2) mmxgen_s(genes_to_genes, genes_to_genes, maskedmxkernel, maxmin1);
    else mxgen_s(genes_to_genes, genes_to_genes, mxkernel, maxmin1);
    if (tfonly) {
      if (transition == 2 || transition == 1) mmxgen(tfs_to_genes, maxmin1, maskedmxkernel, maxmin2);
      else mxgen(tfs_to_genes, maxmin1, mxkernel, maxmin2);
      maxmin_final = maxmin2;
    } else {
      mxgen(genes_to_genes, maxmin1, mxkernel, maxmin2);
      // max(A, A^T) (mcp.cpp:463-474): a no-op for a symmetric MI matrix
      MatrixType maxmin2t = maxmin2.local_transpose();
      splash::pattern::BinaryOp<MatrixType, MatrixType, mcp::kernel::max_kernel<DataType>, MatrixType> maxgen;
      maxgen(maxmin2, maxmin2t, mcp::kernel::max_kernel<DataType>(), maxmin_final);
    }
  } else if (compute == MCP4) {
    MatrixType maxmin1, maxmin2, maxmin3;
    if (tfonly && transition == 3) mmxgen_s(genes_to_genes, genes_to_genes, maskedmxkernel, maxmin1);
    else mxgen_s(genes_to_genes, genes_to_genes, mxkernel, maxmin1);
    if (tfonly) {
      if (transition == 0) mxgen(tfs_to_genes, genes_to_genes, mxkernel, maxmin2);
      else mmxgen(tfs_to_genes, genes_to_genes, maskedmxkernel, maxmin2);
      if (transition == 2 || transition == 3) mmxgen(maxmin2, maxmin1, maskedmxkernel, maxmin3);
      else mxgen(maxmin2, maxmin1, mxkernel, maxmin3);
    } else {
      // the reference recomputes A (x) A as its second partial (mcp.cpp:564-567); it equals maxmin1
      mxgen_s(maxmin1, maxmin1, mxkernel, maxmin3);
    }
    maxmin_final = maxmin3;
  } else if (compute == BIN_MCP) {
    if (tfonly) throw std::runtime_error("transcription factors are not supported when computing MCP from partials");
    mxgen(input1, input2, mxkernel, maxmin_final);
  } else {
    throw std::runtime_error("--method must be 1 (MCP2), 2 (MCP3), 3 (MCP4) or 5 (from partials)");
  }
  printf("Compute max of row-wise min in %f sec\n", t.seconds());
  if (!mm_out.empty() && !mm_out[0].empty())
    mio::write_matrix(mm_out[0], tfonly ? tf_names : genes, genes, maxmin_final);
  t = tool::Timer();
  tolgen(rows_in, maxmin_final, ratio, output);
  printf("Compute MI Ratio in %f sec\n", t.seconds());
  mio::write_matrix(p.str("--output"), tfonly ? tf_names : genes, genes, output);
  printf("output size: %zu x %zu\n", output.rows(), output.columns());
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  cli::Parser p("mcp");
  tool::common_options(p);
  p.option("-f,--tf-input", "transcription factor file, 1 per line.", true, "");
  p.option("--tf-gene-transition", "transition from tf-tf to gene-gene, in [-1, 3]", true, "-1");
  p.option("-m,--method", "Algorithm: MCP2=1, MCP3=2, MCP4=3, MCP from Partials=5", true, "1");
  p.option("--first", "first input, for computing MCP from partials.", true, "");
  p.option("--second", "second input, for computing MCP from partials.", true, "");
  p.option("--diagonal", "set the input MI matrix diagonal to this value", true, "");
  p.option("--clamped", "output is clamped (no effect, as in the reference)", false);
  p.option("-x,--maxmin-output", "filename for the max-min intermediate values", true, "");
  try {
    p.parse(argc, argv);
    if (p.str("--output").empty()) throw std::runtime_error("no output: give -o FILE");
    tool::Timer t;
    int rc = p.seen("--single") ? run<float>(p) : run<double>(p);
    printf("Total Computation in %f sec\n", t.seconds());
    return rc;
  } catch (const std::exception& e) {
    fprintf(stderr, "mcp: %s\n", e.what());
    return 1;
  }
}
