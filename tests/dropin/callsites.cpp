// Call-site check of the drop-in headers: the stage functions below are written against the
// headers at the include paths of the reference's pipeline code, and instantiate the exact
// pattern / kernel-functor template arguments of its call sites:
//   compute_mi        apps/pipeline/mcpnet_helper.hpp:56-99
//   compute_ap_mi     apps/pipeline/mcpnet_helper.hpp:102-167 (add_noise, Rank, histogram, AP kernel)
//   compute_ratio     apps/pipeline/mcpnet_helper.hpp:172-184
//   compute_maxmin1/2/3 + compute_mcp2/3/4   apps/pipeline/mcpnet_helper.hpp:188-205, :247-277, :335-364
//   the mi tool's B-spline stages             apps/correlation/mi.cpp:196-251
// The CPU suite compiles this file with -std=c++11 (the reference's standard, CMakeLists.txt:22);
// the GPU suite runs the built program on a seeded matrix and compares every output with the
// oracle.  Usage: callsites <X.f32> <genes> <samples> <outdir>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <string>
#include <vector>

#include "splash/utils/benchmark.hpp"
#include "splash/utils/report.hpp"
#include "splash/ds/aligned_matrix.hpp"
#include "splash/patterns/pattern.hpp"

#include "mcp/correlation/BSplineMI.hpp"
#include "mcp/correlation/AdaptivePartitioningMI.hpp"
#include "splash/transform/rank.hpp"

#include "mcp/filter/mcp.hpp"
#include "mcp/filter/dpi.hpp"
#include "mcp/transform/stouffer.hpp"
#include "mcp/transform/combine.hpp"

template <typename T, template <typename> class MatrixType>
void stage_mi(MatrixType<T> const& input, int const& bins, MatrixType<T>& dmi) {
  splash::pattern::GlobalTransform<MatrixType<T>, mcp::kernel::MinMaxScale<T>, MatrixType<T>> scale;
  MatrixType<T> scaled(input.rows(), input.columns());
  scale(input, mcp::kernel::MinMaxScale<T>(), scaled);

  const int S = static_cast<int>(input.columns());
  splash::pattern::Transform<MatrixType<T>, mcp::kernel::BSplineWeightsKernel<T>, MatrixType<T>> weigh;
  MatrixType<T> w(scaled.rows(), scaled.columns() * bins + 1);
  weigh(scaled, mcp::kernel::BSplineWeightsKernel<T>(bins, 3, S), w);

  auto t0 = getSysTime();
  MatrixType<T> all = w.allgather();
  FMT_ROOT_PRINT("weights allgather {} s\n", get_duration_s(t0, getSysTime()));

  splash::pattern::InnerProduct<MatrixType<T>, mcp::kernel::BSplineMIKernel<T>, MatrixType<T>, true> mi;
  mi(all, all, mcp::kernel::BSplineMIKernel<T>(bins, S), dmi);
}

template <typename T, template <typename> class MatrixType>
void stage_ap_mi(MatrixType<T> const& input, MatrixType<T>& dmi, bool const& jitter) {
  MatrixType<T> noisy(input.rows(), input.columns());
  if (jitter) {
    using Noise = ::mcp::kernel::add_noise<T, std::uniform_real_distribution<T>, std::default_random_engine>;
    splash::kernel::random_number_generator<> gen;
    Noise noise(gen, 0.0, 0.00001);
    noise(input, noisy);
  } else {
    noisy = input;
  }
  using Ranker = splash::kernel::Rank<T, size_t, 0, false>;
  splash::pattern::Transform<MatrixType<T>, Ranker, splash::ds::aligned_matrix<size_t>> rank;
  splash::ds::aligned_matrix<size_t> ranked(noisy.rows(), noisy.columns());
  rank(noisy, Ranker(), ranked);

  using Hist = mcp::kernel::IntegralCumulativeHistogramKernel<size_t, 0>;
  splash::pattern::Transform<splash::ds::aligned_matrix<size_t>, Hist, splash::ds::aligned_matrix<size_t>> hist;
  splash::ds::aligned_matrix<size_t> histos(ranked.rows(), ranked.columns() + 1);
  hist(ranked, Hist(), histos);

  using APKernel = mcp::kernel::AdaptivePartitionRankMIKernel2<size_t, T>;
  splash::pattern::InnerProduct<splash::ds::aligned_matrix<size_t>, APKernel, MatrixType<T>, true> ap;
  APKernel op(histos.data(), histos.column_bytes());
  ap(ranked, ranked, op, dmi);
}

template <typename T, template <typename> class MatrixType>
void stage_ratio(MatrixType<T> const& mi, MatrixType<T> const& mm, MatrixType<T>& out) {
  splash::utils::partition<size_t> whole = splash::utils::partition<size_t>::make_partition(mm.rows());
  splash::utils::partition<size_t> local(0, mm.rows(), 0);
  out.resize(mm.rows(), mm.columns());
  ::splash::pattern::BinaryOp<MatrixType<T>, MatrixType<T>, mcp::kernel::ratio_kernel<T, T>, MatrixType<T>> op;
  op(mi, whole, mm, local, mcp::kernel::ratio_kernel<T, T>(), out, local);
}

// order 1: A(x)A, 2: A(x)(A(x)A), 3: (A(x)A)(x)(A(x)A)
template <typename T, template <typename> class MatrixType>
void stage_maxmin(MatrixType<T> const& A, int order, MatrixType<T>& out) {
  using MX = mcp::kernel::mcp2_maxmin_kernel<T, false, true>;
  ::splash::pattern::InnerProduct<MatrixType<T>, MX, MatrixType<T>, false> gen;
  ::splash::pattern::InnerProduct<MatrixType<T>, MX, MatrixType<T>, true> gen_s;
  MX k;
  if (order == 1) {
    out.resize(A.rows(), A.columns());
    gen_s(A, A, k, out);
    return;
  }
  MatrixType<T> d1(A.rows(), A.columns());
  gen_s(A, A, k, d1);
  MatrixType<T> m1 = d1.allgather();
  out.resize(A.rows(), m1.columns());
  if (order == 2)
    gen(A, m1, k, out);
  else
    gen_s(m1, m1, k, out);
}

template <typename T, template <typename> class MatrixType>
void stage_mcp(MatrixType<T> const& A, int k, MatrixType<T>& mm, MatrixType<T>& out) {
  stage_maxmin<T, MatrixType>(A, k - 1, mm);
  stage_ratio<T, MatrixType>(A, mm, out);
}

template <typename T>
static void write_raw(const std::string& path, const splash::ds::aligned_matrix<T>& m) {
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) {
    std::perror(path.c_str());
    std::exit(1);
  }
  for (size_t i = 0; i < m.rows(); ++i) std::fwrite(m.data(i), sizeof(T), m.columns(), f);
  std::fclose(f);
}

template <typename T>
static void run_mcp(const splash::ds::aligned_matrix<T>& A, const std::string& dir, const char* tag) {
  for (int k = 2; k <= 4; ++k) {
    splash::ds::aligned_matrix<T> mm, out;
    stage_mcp<T, splash::ds::aligned_matrix>(A, k, mm, out);
    write_raw(dir + "/mm" + std::to_string(k) + tag, mm);
    write_raw(dir + "/mcp" + std::to_string(k) + tag, out);
  }
}

int main(int argc, char** argv) {
  if (argc != 5) {
    std::fprintf(stderr, "usage: %s X.f32 genes samples outdir\n", argv[0]);
    return 2;
  }
  const size_t G = std::strtoull(argv[2], nullptr, 10), S = std::strtoull(argv[3], nullptr, 10);
  const std::string dir = argv[4];
  splash::ds::aligned_matrix<float> X(G, S);
  FILE* f = std::fopen(argv[1], "rb");
  if (!f) {
    std::perror(argv[1]);
    return 1;
  }
  for (size_t i = 0; i < G; ++i)
    if (std::fread(X.data(i), sizeof(float), S, f) != S) {
      std::fprintf(stderr, "short read\n");
      return 1;
    }
  std::fclose(f);
  try {
    splash::ds::aligned_matrix<float> mi(G, G), ap(G, G);
    stage_mi<float, splash::ds::aligned_matrix>(X, 10, mi);
    write_raw(dir + "/mi.f32", mi);
    stage_ap_mi<float, splash::ds::aligned_matrix>(X, ap, true);
    write_raw(dir + "/ap.f32", ap);
    for (size_t i = 0; i < G; ++i) mi.at(i, i) = 0.0f;   // mcpnet's default diagonal (mcpnet.cpp:299-308)
    run_mcp<float>(mi, dir, ".f32");
    splash::ds::aligned_matrix<double> mi64(G, G);
    for (size_t i = 0; i < G; ++i)
      for (size_t j = 0; j < G; ++j) mi64.at(i, j) = mi.at(i, j);
    run_mcp<double>(mi64, dir, ".f64");
  } catch (const mcpgpu::error& e) {
    std::fprintf(stderr, "callsites: %s\n", e.what());
    return 3;
  }
  FMT_ROOT_PRINT("callsites: {} genes x {} samples done\n", G, S);
  FMT_FLUSH();
  return 0;
}
