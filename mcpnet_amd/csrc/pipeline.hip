// The mcpnet stage chain in one call (apps/pipeline/mcpnet.cpp:286-515): MI -> diagonal ->
// TF rows -> MCP2/3/4, with every intermediate in device memory.  The stages are the library's
// own entry points called on device pointers; this file adds the glue kernels (type
// conversion, diagonal, row gather) and owns the intermediates, which live in allocations of
// their own because each stage reuses the handle's workspace.
#include <algorithm>
#include <type_traits>
#include <vector>

#include "internal.h"

namespace {

using mcpg::ceil_div;

int grid_of(int64_t total) { return (int)std::min<int64_t>(std::max<int64_t>(ceil_div(total, 256), 1), 65536); }

// out = (O)in elementwise over an R x C block
template <typename I, typename O>
__global__ void convert_kernel(const I* __restrict__ in, int64_t ldi, int64_t R, int64_t C, O* __restrict__ out,
                               int64_t ldo) {
  const int64_t total = R * C;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = t / C, c = t % C;
    out[r * ldo + c] = static_cast<O>(in[r * ldi + c]);
  }
}

// mcpnet.cpp:299-308: A[i][i] = v
template <typename T>
__global__ void set_diagonal_kernel(T* __restrict__ A, int64_t lda, int64_t n, T v) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    A[i * lda + i] = v;
}

// filter_local_mat_rows_by_tfs (mcpnet_helper.hpp:1147-1170): out[r] = A[rows[r]]
template <typename T>
__global__ void gather_rows_kernel(const T* __restrict__ A, int64_t lda, const int32_t* __restrict__ rows, int64_t R,
                                   int64_t C, T* __restrict__ out, int64_t ldo) {
  const int64_t total = R * C;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total; t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = t / C, c = t % C;
    out[r * ldo + c] = A[(int64_t)rows[r] * lda + c];
  }
}

// One device allocation owned by the call (freed on every return path).
struct Buf {
  void* p = nullptr;
  Buf() = default;
  Buf(const Buf&) = delete;
  Buf& operator=(const Buf&) = delete;
  ~Buf() {
    if (p) (void)hipFree(p);
  }
};

template <typename T>
int alloc(mcpg_ctx* h, Buf& b, size_t count, T** out) {
  MCPG_HIP(h, hipMalloc(&b.p, std::max<size_t>(count * sizeof(T), 256)));
  *out = static_cast<T*>(b.p);
  return MCPG_OK;
}

// Marks the stages as parts of this call (see mcpg_ctx::nest).
struct Nest {
  mcpg_ctx* h;
  explicit Nest(mcpg_ctx* c) : h(c) { ++h->nest; }
  ~Nest() { --h->nest; }
};

template <typename T>
struct Stage;
template <>
struct Stage<float> {
  static int maxmin(mcpg_ctx* h, const float* A, int64_t lda, const float* B, int64_t ldb, const float* km, int64_t M,
                    int64_t N, int64_t K, int sym, float* C, int64_t ldc) {
    return mcpg_maxmin_f32(h, MCPG_DEVICE, A, lda, B, ldb, km, M, N, K, sym, C, ldc);
  }
  static int ratio(mcpg_ctx* h, const float* A, int64_t lda, const float* Mm, int64_t ldm, int64_t M, int64_t N,
                   float* out, int64_t ldo) {
    return mcpg_ratio_f32(h, MCPG_DEVICE, A, lda, Mm, ldm, M, N, out, ldo);
  }
  static int mcp(mcpg_ctx* h, const float* A, int64_t lda, int64_t N, int k, float* out, int64_t ldo) {
    return mcpg_mcp_f32(h, MCPG_DEVICE, A, lda, N, k, nullptr, 0, out, ldo);
  }
};
template <>
struct Stage<double> {
  static int maxmin(mcpg_ctx* h, const double* A, int64_t lda, const double* B, int64_t ldb, const double* km,
                    int64_t M, int64_t N, int64_t K, int sym, double* C, int64_t ldc) {
    return mcpg_maxmin_f64(h, MCPG_DEVICE, A, lda, B, ldb, km, M, N, K, sym, C, ldc);
  }
  static int ratio(mcpg_ctx* h, const double* A, int64_t lda, const double* Mm, int64_t ldm, int64_t M, int64_t N,
                   double* out, int64_t ldo) {
    return mcpg_ratio_f64(h, MCPG_DEVICE, A, lda, Mm, ldm, M, N, out, ldo);
  }
  static int mcp(mcpg_ctx* h, const double* A, int64_t lda, int64_t N, int k, double* out, int64_t ldo) {
    return mcpg_mcp_f64(h, MCPG_DEVICE, A, lda, N, k, nullptr, 0, out, ldo);
  }
};

// The G x G MI matrix (FP32-class arithmetic) of the G x S expression matrix X in device memory.
// With MI64 (adaptive partitioning only) the MI is written there in double instead.
int mi_f32(mcpg_ctx* h, const float* X, int64_t G, int64_t S, int64_t ldx, int method, int bins, int order,
           int precision, uint64_t seed, float* MI, double* MI64 = nullptr) {
  if (method == 0) return mcpg_bspline_mi_f32(h, MCPG_DEVICE, X, G, S, ldx, bins, order, precision, MI, G);
  // compute_ap_mi (mcpnet_helper.hpp:102-167): noise, descending ranks, AP-MI on the ranks
  Buf by, br;
  float* y;
  int32_t* r;
  int st;
  if ((st = alloc(h, by, (size_t)G * S, &y)) || (st = alloc(h, br, (size_t)G * S, &r))) return st;
  if ((st = mcpg_add_noise_f32(h, MCPG_DEVICE, X, G, S, ldx, seed, y, S))) return st;
  if ((st = mcpg_rank_f32(h, MCPG_DEVICE, y, G, S, S, r, S))) return st;
  if (MI64) return mcpg_ap_mi_f64(h, MCPG_DEVICE, r, G, S, S, MI64, G);
  return mcpg_ap_mi_i32(h, MCPG_DEVICE, r, G, S, S, MI, G);
}

template <typename T>
int mcpnet_entry(mcpg_ctx* h, int ptrs, const T* X, int64_t G, int64_t S, int64_t ldx, int mi_method, int bins,
                 int order, int precision, uint64_t seed, double diagonal, const T* tfs, T* mi_out, int64_t ldmi,
                 T* mcp2, int64_t ld2, T* mcp3, int64_t ld3, T* mcp4, int64_t ld4) {
  MCPG_CHECK_HANDLE(h);
  mcpg::begin_call(h);
  if (G < 0 || S < 0) return mcpg::set_error(h, MCPG_ERR_SHAPE, "negative dimension");
  if (mi_method < -1 || mi_method > 1) return mcpg::set_error(h, MCPG_ERR_INVALID_ARG, "mi_method must be -1, 0 or 1");
  if (ptrs != MCPG_HOST && ptrs != MCPG_DEVICE) return mcpg::set_error(h, MCPG_ERR_INVALID_ARG, "bad ptrs kind");
  if (G == 0) return MCPG_OK;
  if (!X) return mcpg::set_error(h, MCPG_ERR_INVALID_ARG, "null input matrix");
  const int64_t cols = mi_method < 0 ? G : S;
  if (ldx < cols) return mcpg::set_error(h, MCPG_ERR_INVALID_ARG, "leading dimension too small");
  T* outs[4] = {mi_out, mcp2, mcp3, mcp4};
  const int64_t lds[4] = {ldmi, ld2, ld3, ld4};
  for (int q = 0; q < 4; ++q)
    if (outs[q] && lds[q] < G) return mcpg::set_error(h, MCPG_ERR_INVALID_ARG, "leading dimension too small");
  MCPG_HIP(h, hipSetDevice(h->device));
  Nest nest(h);
  constexpr bool f32 = std::is_same<T, float>::value;
  const bool host = ptrs == MCPG_HOST;
  const hipMemcpyKind in_kind = host ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice;
  const hipMemcpyKind out_kind = host ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
  int st;

  // TF rows: the caller's mask values are used as the masked kernel's kmask (mcpnet_helper.hpp's
  // load_tfs stores numeric_limits<double>::max() / lowest() in T)
  std::vector<T> htf;
  std::vector<int32_t> rows;
  if (tfs) {
    htf.resize(G);
    if (host) std::copy(tfs, tfs + G, htf.begin());
    else MCPG_HIP(h, hipMemcpy(htf.data(), tfs, G * sizeof(T), hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < G; ++i)
      if (htf[i] > (T)0) rows.push_back((int32_t)i);
    if (rows.empty()) return mcpg::set_error(h, MCPG_ERR_INVALID_ARG, "the TF list selects no gene");
  }
  const int64_t R = tfs ? (int64_t)rows.size() : G;

  // ---- MI (G x G, T) in device memory
  Buf bmi;
  T* dmi;
  if ((st = alloc(h, bmi, (size_t)G * G, &dmi))) return st;
  if (mi_method < 0) {
    if ((st = mcpg::copy_2d(h, dmi, G * sizeof(T), X, ldx * sizeof(T), G * sizeof(T), G, in_kind))) return st;
  } else if (mi_method == 0 && precision == MCPG_FP64) {
    // B-spline MI of double rows in double (the reference's double DataType); the FP32 chain
    // rounds it to float
    Buf bxd, bmid;
    const double* xd = reinterpret_cast<const double*>(X);
    int64_t ldxd = ldx;
    if (f32 || host) {
      double* dxd;
      if ((st = alloc(h, bxd, (size_t)G * S, &dxd))) return st;
      if (!f32) {
        if ((st = mcpg::copy_2d(h, dxd, S * 8, X, ldx * 8, S * 8, G, in_kind))) return st;
      } else if (host) {
        std::vector<double> tmp((size_t)G * S);
        for (int64_t i = 0; i < G; ++i)
          for (int64_t j = 0; j < S; ++j) tmp[i * S + j] = static_cast<double>(X[i * ldx + j]);
        MCPG_HIP(h, hipMemcpy(dxd, tmp.data(), tmp.size() * sizeof(double), hipMemcpyHostToDevice));
      } else {
        mcpg::LaunchScope ls(h, "convert_kernel");
        hipLaunchKernelGGL((convert_kernel<T, double>), dim3(grid_of(G * S)), dim3(256), 0, h->stream, X, ldx, G, S,
                           dxd, S);
      }
      xd = dxd;
      ldxd = S;
    }
    double* dmid = reinterpret_cast<double*>(dmi);
    if (f32 && (st = alloc(h, bmid, (size_t)G * G, &dmid))) return st;
    if ((st = mcpg_bspline_mi_f64(h, MCPG_DEVICE, xd, G, S, ldxd, bins, order, dmid, G))) return st;
    if (f32) {
      mcpg::LaunchScope ls(h, "convert_kernel");
      hipLaunchKernelGGL((convert_kernel<double, T>), dim3(grid_of(G * G)), dim3(256), 0, h->stream, dmid, G, G, G,
                         dmi, G);
    }
  } else {
    // the MI kernels take FP32 expression values
    Buf bx, bmif;
    const float* xf;
    int64_t ldxf = S;
    if (f32 && !host) {
      xf = reinterpret_cast<const float*>(X);
      ldxf = ldx;
    } else {
      float* dx;
      if ((st = alloc(h, bx, (size_t)G * S, &dx))) return st;
      if (f32) {
        if ((st = mcpg::copy_2d(h, dx, S * sizeof(float), X, ldx * sizeof(T), S * sizeof(T), G, in_kind))) return st;
      } else if (host) {
        std::vector<float> tmp((size_t)G * S);
        for (int64_t i = 0; i < G; ++i)
          for (int64_t j = 0; j < S; ++j) tmp[i * S + j] = static_cast<float>(X[i * ldx + j]);
        MCPG_HIP(h, hipMemcpy(dx, tmp.data(), tmp.size() * sizeof(float), hipMemcpyHostToDevice));
      } else {
        mcpg::LaunchScope ls(h, "convert_kernel");
        hipLaunchKernelGGL((convert_kernel<T, float>), dim3(grid_of(G * S)), dim3(256), 0, h->stream, X, ldx, G, S, dx,
                           S);
      }
      xf = dx;
    }
    // the double chain keeps the AP-MI in double (AdaptivePartitionRankMIKernel2<IT, double>)
    const bool ap64 = !f32 && mi_method == 1;
    float* dmif = reinterpret_cast<float*>(dmi);
    if (!f32 && !ap64 && (st = alloc(h, bmif, (size_t)G * G, &dmif))) return st;
    if ((st = mi_f32(h, xf, G, S, ldxf, mi_method, bins, order, precision, seed, dmif,
                     ap64 ? reinterpret_cast<double*>(dmi) : nullptr)))
      return st;
    if (!f32 && !ap64) {
      mcpg::LaunchScope ls(h, "convert_kernel");
      hipLaunchKernelGGL((convert_kernel<float, T>), dim3(grid_of(G * G)), dim3(256), 0, h->stream, dmif, G, G, G, dmi,
                         G);
    }
  }
  if (diagonal >= 0.0) {
    mcpg::LaunchScope ls(h, "set_diagonal_kernel");
    hipLaunchKernelGGL(set_diagonal_kernel<T>, dim3(grid_of(G)), dim3(256), 0, h->stream, dmi, G, G, (T)diagonal);
  }

  // ---- TF rows of the MI and the device mask
  Buf bmask, brows, bmit;
  T* dmask = nullptr;
  T* dmit = dmi;
  if (tfs) {
    int32_t* drows;
    if ((st = alloc(h, bmask, (size_t)G, &dmask)) || (st = alloc(h, brows, rows.size(), &drows)) ||
        (st = alloc(h, bmit, (size_t)R * G, &dmit)))
      return st;
    MCPG_HIP(h, hipMemcpyAsync(dmask, htf.data(), G * sizeof(T), hipMemcpyHostToDevice, h->stream));
    MCPG_HIP(h, hipMemcpyAsync(drows, rows.data(), rows.size() * sizeof(int32_t), hipMemcpyHostToDevice, h->stream));
    mcpg::LaunchScope ls(h, "gather_rows_kernel");
    hipLaunchKernelGGL(gather_rows_kernel<T>, dim3(grid_of(R * G)), dim3(256), 0, h->stream, dmi, G, drows, R, G, dmit,
                       G);
  }
  // the MI download overlaps the max-min products (host outputs: on the side stream)
  mcpg::Download mi_dl;
  if (mi_out && host && (st = mi_dl.start(h, mi_out, ldmi * sizeof(T), dmi, G * sizeof(T), G * sizeof(T), G)))
    return st;
  if (mi_out && !host && (st = mcpg::copy_2d(h, mi_out, ldmi * sizeof(T), dmi, G * sizeof(T), G * sizeof(T), G, out_kind)))
    return st;

  // ---- MCP2/3/4 (compute_mcp{2,3,4}, or compute_mcp{2,3,4}_tfs with tf_gene_transition k - 1)
  // With host outputs each order's scores (or row panel of them) come back on the side stream
  // while the next product (or panel) is computed; the downloads run one after another (they share
  // the side stream's pinned buffers), and each order's device copy lives until the end of the call.
  Buf bm1;
  T* dm1 = nullptr;   // masked A (x) A (gene-TF-gene), shared by the TF orders 3 and 4
  Buf bouts[3], btmp[3][2];
  constexpr int kMaxDownloads = 8;   // two orders whole, the last in at most four row panels
  mcpg::Download dls[kMaxDownloads];
  int ndl = 0;
  auto download = [&](T* dst, int64_t ld, const T* src, int64_t rows) -> int {
    int rc;
    if (ndl == kMaxDownloads) return mcpg::set_error(h, MCPG_ERR_UNSUPPORTED, "too many chain downloads");
    if ((rc = mi_dl.wait(h)) || (ndl > 0 && (rc = dls[ndl - 1].wait(h)))) return rc;
    return dls[ndl++].start(h, dst, ld * sizeof(T), src, G * sizeof(T), G * sizeof(T), rows);
  };
  // Without TFs and with two or more orders the max-min matrices are shared: M1 = A (x) A once,
  // M2 = A (x) M1 and M3 = M1 (x) M1 as asked (compute_maxmins, mcpnet_helper.hpp:436-460), and
  // MCP_k = ratio(A, M_{k-1}); each compute_mcp{k} alone would form A (x) A again.
  Buf bmx[3];
  T* dmx[3] = {nullptr, nullptr, nullptr};
  const bool shared = !tfs && (mcp2 ? 1 : 0) + (mcp3 ? 1 : 0) + (mcp4 ? 1 : 0) >= 2;
  // Rows [r0, r1) of order k's scores from what is on the stream so far, then their download on
  // the side stream.  Only the shared chain hands on row panels; the other forms whole orders.
  auto emit = [&](int k, int64_t r0, int64_t r1) -> int {
    T* out = outs[k - 1];
    if (!out) return MCPG_OK;
    Buf& bout = bouts[k - 2];
    int st;
    T* dout = out;
    int64_t ldo = lds[k - 1];
    if (host) {   // one device copy per order, filled panel by panel
      if (!bout.p && (st = alloc(h, bout, (size_t)R * G, &dout))) return st;
      dout = static_cast<T*>(bout.p);
      ldo = G;
    }
    if (shared) {
      if ((st = Stage<T>::ratio(h, dmi + r0 * G, G, dmx[k - 2] + r0 * G, G, r1 - r0, G, dout + r0 * ldo, ldo)))
        return st;
    } else if (!tfs) {
      if ((st = Stage<T>::mcp(h, dmi, G, G, k, dout, ldo))) return st;
    } else {
      if (k >= 3 && !dm1) {
        if ((st = alloc(h, bm1, (size_t)G * G, &dm1))) return st;
        if ((st = Stage<T>::maxmin(h, dmi, G, dmi, G, dmask, G, G, G, 1, dm1, G))) return st;
      }
      Buf& bmm = btmp[k - 2][0];   // kept to the end of the call: a free would wait for the download
      Buf& bm2 = btmp[k - 2][1];
      T* dmm;
      if ((st = alloc(h, bmm, (size_t)R * G, &dmm))) return st;
      if (k == 2) {          // tfs-tfs-genes
        st = Stage<T>::maxmin(h, dmit, G, dmi, G, dmask, R, G, G, 0, dmm, G);
      } else if (k == 3) {   // tf-(tf-tf-gene)
        st = Stage<T>::maxmin(h, dmit, G, dm1, G, dmask, R, G, G, 0, dmm, G);
      } else {               // (tf-tf-gene)-(tf-gene); its first factor is MCP2's max-min
        T* dm2 = mcp2 ? reinterpret_cast<T*>(btmp[0][0].p) : nullptr;
        if (!dm2) {
          if ((st = alloc(h, bm2, (size_t)R * G, &dm2))) return st;
          if ((st = Stage<T>::maxmin(h, dmit, G, dmi, G, dmask, R, G, G, 0, dm2, G))) return st;
        }
        st = Stage<T>::maxmin(h, dm2, G, dm1, G, dmask, R, G, G, 0, dmm, G);
      }
      if (st || (st = Stage<T>::ratio(h, dmit, G, dmm, G, R, G, dout, ldo))) return st;
    }
    if (host && (st = download(out + r0 * lds[k - 1], lds[k - 1], dout + r0 * ldo, r1 - r0))) return st;
    return mcpg::finish(h, "mcpg_mcpnet");
  };
  if (shared) {
    for (int q = 0; q < 3; ++q)
      if ((q == 0 || outs[q + 1]) && (st = alloc(h, bmx[q], (size_t)G * G, &dmx[q]))) return st;
    // host outputs: order k is emitted as soon as M_{k-1} (or a row panel of the last product) is
    // on the stream, so that its download overlaps the next product or panel; device outputs have
    // nothing to overlap, and the ratios follow the whole products
    const mcpg::ChainHook after = [&](int q, int64_t r0, int64_t r1) { return emit(q + 2, r0, r1); };
    if ((st = mcpg::maxmins_chain(h, dmi, G, dmx[0], dmx[1], dmx[2], host ? &after : nullptr))) return st;
    for (int k = 2; !host && k <= 4; ++k)
      if ((st = emit(k, 0, G))) return st;
  } else {
    for (int k = 2; k <= 4; ++k)
      if ((st = emit(k, 0, R))) return st;
  }
  if ((st = mi_dl.wait(h))) return st;
  for (int q = 0; q < ndl; ++q)
    if ((st = dls[q].wait(h))) return st;
  return mcpg::finish(h, "mcpg_mcpnet");
}

}  // namespace

extern "C" {

int mcpg_mcpnet_f32(mcpg_handle h, int ptrs, const float* X, int64_t G, int64_t S, int64_t ldx, int mi_method, int bins,
                    int order, int precision, uint64_t noise_seed, double diagonal, const float* tfs, float* mi,
                    int64_t ldmi, float* mcp2, int64_t ld2, float* mcp3, int64_t ld3, float* mcp4, int64_t ld4) {
  return mcpnet_entry<float>(h, ptrs, X, G, S, ldx, mi_method, bins, order, precision, noise_seed, diagonal, tfs, mi,
                             ldmi, mcp2, ld2, mcp3, ld3, mcp4, ld4);
}

int mcpg_mcpnet_f64(mcpg_handle h, int ptrs, const double* X, int64_t G, int64_t S, int64_t ldx, int mi_method,
                    int bins, int order, int precision, uint64_t noise_seed, double diagonal, const double* tfs,
                    double* mi, int64_t ldmi, double* mcp2, int64_t ld2, double* mcp3, int64_t ld3, double* mcp4,
                    int64_t ld4) {
  return mcpnet_entry<double>(h, ptrs, X, G, S, ldx, mi_method, bins, order, precision, noise_seed, diagonal, tfs, mi,
                              ldmi, mcp2, ld2, mcp3, ld3, mcp4, ld4);
}

}  // extern "C"
