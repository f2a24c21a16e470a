// Handle lifecycle, error reporting and workspace management of libmcpgpu.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <thread>
#include <vector>

#include "internal.h"

namespace mcpg {

int set_error(mcpg_ctx* h, int status, const std::string& msg) {
  if (h) h->last_error = msg;
  return status;
}

int hip_error(mcpg_ctx* h, hipError_t e, const char* where) {
  std::string msg = std::string(where) + ": " + hipGetErrorString(e);
  int status = (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) ? MCPG_ERR_OUT_OF_MEMORY : MCPG_ERR_HIP;
  if (e == hipErrorNoDevice || e == hipErrorInvalidDevice) status = MCPG_ERR_NO_DEVICE;
  (void)hipGetLastError();  // clear the sticky-free error so later calls on the handle start clean
  return set_error(h, status, msg);
}

void* workspace(mcpg_ctx* h, size_t bytes, int* status) {
  *status = MCPG_OK;
  if (bytes <= h->ws_bytes) return h->ws;
  if (h->ws) {
    hipError_t e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) { *status = hip_error(h, e, "hipStreamSynchronize"); return nullptr; }
    (void)hipFree(h->ws);
    h->ws = nullptr;
    h->ws_bytes = 0;
  }
  size_t want = align_up(bytes, size_t(1) << 20);
  hipError_t e = hipMalloc(&h->ws, want);
  if (e != hipSuccess) {
    h->ws = nullptr;
    *status = hip_error(h, e, "hipMalloc(workspace)");
    char buf[96];
    snprintf(buf, sizeof buf, " (requested %zu bytes)", want);
    h->last_error += buf;
    return nullptr;
  }
  h->ws_bytes = want;
  return h->ws;
}

// Pageable host memory moves far below the link rate when the runtime stages it itself.  Large
// copies therefore go through two pinned buffers: the CPU packs (or unpacks) one chunk of rows,
// split over a few threads, while the DMA engine moves the other.
constexpr size_t kPinBytes = size_t(32) << 20;
constexpr size_t kPinMin = size_t(8) << 20;

static void copy_rows_parallel(char* dst, size_t dpitch, const char* src, size_t spitch, size_t width, size_t rows) {
  const size_t bytes = width * rows;
  const int nt = (int)std::min<size_t>(8, std::max<size_t>(1, bytes >> 22));   // one thread per 4 MB, at most 8
  auto part = [&](int t) {
    const size_t r0 = rows * t / nt, r1 = rows * (t + 1) / nt;
    if (dpitch == width && spitch == width) {
      std::memcpy(dst + r0 * width, src + r0 * width, (r1 - r0) * width);
    } else {
      for (size_t r = r0; r < r1; ++r) std::memcpy(dst + r * dpitch, src + r * spitch, width);
    }
  };
  if (nt == 1) { part(0); return; }
  std::vector<std::thread> th;
  for (int t = 1; t < nt; ++t) th.emplace_back(part, t);
  part(0);
  for (auto& x : th) x.join();
}

static int ensure_pinned(mcpg_ctx* h) {
  for (int b = 0; b < 2; ++b) {
    if (!h->pin[b]) MCPG_HIP(h, hipHostMalloc(&h->pin[b], kPinBytes, hipHostMallocDefault));
    if (!h->pin_ev[b]) MCPG_HIP(h, hipEventCreateWithFlags(&h->pin_ev[b], hipEventDisableTiming));
  }
  return MCPG_OK;
}

void* pinned_staging(mcpg_ctx* h, int b, size_t bytes, int* st) {
  *st = MCPG_OK;
  if (bytes > kPinBytes) return nullptr;
  if ((*st = ensure_pinned(h))) return nullptr;
  if (h->pin_busy[b]) {
    const hipError_t e = hipEventSynchronize(h->pin_ev[b]);
    if (e != hipSuccess) {
      *st = hip_error(h, e, "hipEventSynchronize");
      return nullptr;
    }
    h->pin_busy[b] = false;
  }
  return h->pin[b];
}

int pinned_upload(mcpg_ctx* h, int b, void* dst, size_t bytes) {
  MCPG_HIP(h, hipMemcpyAsync(dst, h->pin[b], bytes, hipMemcpyHostToDevice, h->stream));
  MCPG_HIP(h, hipEventRecord(h->pin_ev[b], h->stream));
  h->pin_busy[b] = true;
  return MCPG_OK;
}

int Download::start(mcpg_ctx* h, void* dst, size_t dpitch, const void* src, size_t spitch, size_t width_bytes,
                    size_t rows) {
  if (rows == 0 || width_bytes == 0) return MCPG_OK;
  if (width_bytes * rows < kPinMin || width_bytes > kPinBytes)
    return copy_2d(h, dst, dpitch, src, spitch, width_bytes, rows, hipMemcpyDeviceToHost);
  if (!h->side) MCPG_HIP(h, hipStreamCreateWithFlags(&h->side, hipStreamNonBlocking));
  for (int b = 0; b < 2; ++b)
    if (!h->side_pin[b]) MCPG_HIP(h, hipHostMalloc(&h->side_pin[b], kPinBytes, hipHostMallocDefault));
  for (int b = 0; b < 3; ++b)
    if (!h->side_ev[b]) MCPG_HIP(h, hipEventCreateWithFlags(&h->side_ev[b], hipEventDisableTiming));
  MCPG_HIP(h, hipEventRecord(h->side_ev[2], h->stream));
  MCPG_HIP(h, hipStreamWaitEvent(h->side, h->side_ev[2], 0));
  const size_t per = kPinBytes / width_bytes, nchunks = (rows + per - 1) / per;
  char* d = static_cast<char*>(dst);
  const char* s = static_cast<const char*>(src);
  th = std::thread([=]() {
    // a new host thread starts on device 0: make the handle's device current
    if (hipSetDevice(h->device) != hipSuccess) {
      status = MCPG_ERR_HIP;
      (void)hipStreamSynchronize(h->side);
      return;
    }
    auto enqueue = [&](size_t c) -> hipError_t {
      const int b = (int)(c & 1);
      const size_t r0 = c * per, n = std::min(per, rows - r0);
      hipError_t e = hipMemcpy2DAsync(h->side_pin[b], width_bytes, s + r0 * spitch, spitch, width_bytes, n,
                                      hipMemcpyDeviceToHost, h->side);
      return e != hipSuccess ? e : hipEventRecord(h->side_ev[b], h->side);
    };
    hipError_t e = enqueue(0);
    for (size_t c = 0; e == hipSuccess && c < nchunks; ++c) {
      if (c + 1 < nchunks && (e = enqueue(c + 1)) != hipSuccess) break;
      const int b = (int)(c & 1);
      const size_t r0 = c * per, n = std::min(per, rows - r0);
      if ((e = hipEventSynchronize(h->side_ev[b])) != hipSuccess) break;
      copy_rows_parallel(d + r0 * dpitch, dpitch, static_cast<const char*>(h->side_pin[b]), width_bytes, width_bytes, n);
    }
    // nothing queued on the side stream outlives the thread, also after an error
    const hipError_t es = hipStreamSynchronize(h->side);
    if (e != hipSuccess || es != hipSuccess) status = MCPG_ERR_HIP;
  });
  return MCPG_OK;
}

int Download::wait(mcpg_ctx* h) {
  if (th.joinable()) th.join();
  if (status) return set_error(h, status, "side-stream download failed");
  return MCPG_OK;
}

int copy_2d(mcpg_ctx* h, void* dst, size_t dpitch, const void* src, size_t spitch, size_t width_bytes,
            size_t rows, hipMemcpyKind kind) {
  if (rows == 0 || width_bytes == 0) return MCPG_OK;
  const bool h2d = kind == hipMemcpyHostToDevice, d2h = kind == hipMemcpyDeviceToHost;
  if ((!h2d && !d2h) || width_bytes * rows < kPinMin || width_bytes > kPinBytes) {
    MCPG_HIP(h, hipMemcpy2DAsync(dst, dpitch, src, spitch, width_bytes, rows, kind, h->stream));
    return MCPG_OK;
  }
  int st;
  if ((st = ensure_pinned(h))) return st;
  const size_t per = kPinBytes / width_bytes;          // rows per chunk
  const size_t nchunks = (rows + per - 1) / per;
  char* d = static_cast<char*>(dst);
  const char* s = static_cast<const char*>(src);
  if (h2d) {
    for (size_t c = 0; c < nchunks; ++c) {
      const int b = (int)(c & 1);
      const size_t r0 = c * per, n = std::min(per, rows - r0);
      if (h->pin_busy[b]) MCPG_HIP(h, hipEventSynchronize(h->pin_ev[b]));
      copy_rows_parallel(static_cast<char*>(h->pin[b]), width_bytes, s + r0 * spitch, spitch, width_bytes, n);
      MCPG_HIP(h, hipMemcpy2DAsync(d + r0 * dpitch, dpitch, h->pin[b], width_bytes, width_bytes, n,
                                   hipMemcpyHostToDevice, h->stream));
      MCPG_HIP(h, hipEventRecord(h->pin_ev[b], h->stream));
      h->pin_busy[b] = true;
    }
    return MCPG_OK;
  }
  // device -> host: chunk c + 1 is in flight while the CPU unpacks chunk c
  auto enqueue = [&](size_t c) -> int {
    const int b = (int)(c & 1);
    const size_t r0 = c * per, n = std::min(per, rows - r0);
    MCPG_HIP(h, hipMemcpy2DAsync(h->pin[b], width_bytes, s + r0 * spitch, spitch, width_bytes, n,
                                 hipMemcpyDeviceToHost, h->stream));
    MCPG_HIP(h, hipEventRecord(h->pin_ev[b], h->stream));
    h->pin_busy[b] = true;
    return MCPG_OK;
  };
  if ((st = enqueue(0))) return st;
  for (size_t c = 0; c < nchunks; ++c) {
    if (c + 1 < nchunks && (st = enqueue(c + 1))) return st;
    const int b = (int)(c & 1);
    const size_t r0 = c * per, n = std::min(per, rows - r0);
    MCPG_HIP(h, hipEventSynchronize(h->pin_ev[b]));
    copy_rows_parallel(d + r0 * dpitch, dpitch, static_cast<const char*>(h->pin[b]), width_bytes, width_bytes, n);
    h->pin_busy[b] = false;
  }
  return MCPG_OK;
}

int reset_flags(mcpg_ctx* h) {
  MCPG_HIP(h, hipMemsetAsync(h->d_flags, 0, kFlagWords * sizeof(int), h->stream));
  return MCPG_OK;
}

int read_flags(mcpg_ctx* h, int* flags, int n) {
  MCPG_HIP(h, hipMemcpyAsync(flags, h->d_flags, n * sizeof(int), hipMemcpyDeviceToHost, h->stream));
  MCPG_HIP(h, hipStreamSynchronize(h->stream));
  return MCPG_OK;
}

int finish(mcpg_ctx* h, const char* where) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_error(h, e, where);
  e = hipStreamSynchronize(h->stream);
  if (e != hipSuccess) return hip_error(h, e, where);
  return MCPG_OK;
}

const uint32_t* tile_order(mcpg_ctx* h, int64_t ni, int64_t nj, int ratio, int gi, int gj, int xcd_deal,
                           int* status) {
  *status = MCPG_OK;
  if (ni > 65535 || nj > 65535) {
    *status = set_error(h, MCPG_ERR_UNSUPPORTED, "tile grid too large for the packed tile order");
    return nullptr;
  }
  gi = std::max(gi, 1);
  gj = std::max(gj, 1);
  // layout of the table in launch order: with xcd_deal, tiles are dealt so that each XCD (blocks
  // go to the 8 XCDs round-robin by id) receives one contiguous run of the blocked order
  const int xcd = xcd_deal != 0;
  for (const auto& o : h->orders)
    if (o.ni == ni && o.nj == nj && o.ratio == ratio && o.gi == gi && o.gj == gj && o.xcd == xcd) return o.d;
  std::vector<uint32_t> blocked;
  for (int64_t bi = 0; bi < ni; bi += gi)
    for (int64_t bj = 0; bj < nj; bj += gj)
      for (int64_t ti = bi; ti < std::min(ni, bi + gi); ++ti)
        for (int64_t tj = std::max(bj, ratio * ti); tj < std::min(nj, bj + gj); ++tj)
          blocked.push_back((uint32_t)((ti << 16) | tj));
  std::vector<uint32_t> order(blocked.size());
  const int64_t total = (int64_t)blocked.size();
  for (int64_t b = 0; b < total; ++b) {
    int64_t l = b;
    if (xcd) {
      const int64_t x = b % 8, local = b / 8, q = total / 8, r = total % 8;
      l = x * q + (x < r ? x : r) + local;
    }
    order[b] = blocked[l];
  }
  if (h->orders.size() >= 4) {  // keep a few grids (MI and max-min alternate in a pipeline)
    hipError_t e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) { *status = hip_error(h, e, "hipStreamSynchronize"); return nullptr; }
    (void)hipFree(h->orders.front().d);
    h->orders.erase(h->orders.begin());
  }
  uint32_t* d = nullptr;
  hipError_t e = hipMalloc(&d, std::max<size_t>(order.size(), 1) * sizeof(uint32_t));
  if (e != hipSuccess) { *status = hip_error(h, e, "hipMalloc(tile order)"); return nullptr; }
  e = hipMemcpy(d, order.data(), order.size() * sizeof(uint32_t), hipMemcpyHostToDevice);
  if (e != hipSuccess) { (void)hipFree(d); *status = hip_error(h, e, "hipMemcpy(tile order)"); return nullptr; }
  mcpg_ctx::TileOrder o;
  o.d = d; o.ni = ni; o.nj = nj; o.ratio = ratio; o.gi = gi; o.gj = gj; o.xcd = xcd; o.count = total;
  h->orders.push_back(o);
  return d;
}

const uint32_t* tri_order(mcpg_ctx* h, int64_t nb, int group, int xcd_deal, int* status) {
  return tile_order(h, nb, nb, 1, group, group, xcd_deal, status);
}

int env_group(const char* name, int def) {
  const char* v = getenv(name);
  if (!v) return def;
  int g = atoi(v);
  return (g >= 0 && g <= 256) ? g : def;
}

void begin_call(mcpg_ctx* h) {
  if (h->nest > 0) return;
  h->last_launches = 0;
  h->last_kernel = "";
  h->timings.clear();
  h->events_used = 0;
}

static hipEvent_t pooled_event(mcpg_ctx* h) {
  if (h->events_used == h->event_pool.size()) {
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    h->event_pool.push_back(e);
  }
  return h->event_pool[h->events_used++];
}

LaunchScope::LaunchScope(mcpg_ctx* hh, const char* name) : h(hh) {
  h->last_launches++;
  h->last_kernel = name;
  if (!h->profiling) return;
  hipEvent_t start = pooled_event(h);
  stop = pooled_event(h);
  if (!start || !stop) { stop = nullptr; return; }
  (void)hipEventRecord(start, h->stream);
  h->timings.push_back(mcpg_ctx::Timing{name, start, stop});
}

LaunchScope::~LaunchScope() {
  if (stop) (void)hipEventRecord(stop, h->stream);
}

}  // namespace mcpg

extern "C" {

int mcpg_create(mcpg_handle* out, int device) {
  if (!out) return MCPG_ERR_INVALID_ARG;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return MCPG_ERR_NO_DEVICE;
  if (device < 0 || device >= n) return MCPG_ERR_NO_DEVICE;
  mcpg_ctx* h = new (std::nothrow) mcpg_ctx();
  if (!h) return MCPG_ERR_OUT_OF_MEMORY;
  h->device = device;
  if (hipSetDevice(device) != hipSuccess) { delete h; return MCPG_ERR_NO_DEVICE; }
  if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) { delete h; return MCPG_ERR_HIP; }
  if (hipMalloc(&h->d_flags, mcpg::kFlagWords * sizeof(int)) != hipSuccess) {
    (void)hipStreamDestroy(h->stream);
    delete h;
    return MCPG_ERR_OUT_OF_MEMORY;
  }
  mcpg::init_scratch(h);
  *out = h;
  return MCPG_OK;
}

int mcpg_destroy(mcpg_handle h) {
  if (!h) return MCPG_OK;
  (void)hipSetDevice(h->device);
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  for (hipEvent_t e : h->event_pool) (void)hipEventDestroy(e);
  if (h->mm_tail) (void)hipFree(h->mm_tail);
  for (int b = 0; b < 2; ++b) {
    if (h->pin_ev[b]) (void)hipEventDestroy(h->pin_ev[b]);
    if (h->pin[b]) (void)hipHostFree(h->pin[b]);
    if (h->side_pin[b]) (void)hipHostFree(h->side_pin[b]);
  }
  for (hipEvent_t e : h->side_ev)
    if (e) (void)hipEventDestroy(e);
  if (h->side) (void)hipStreamDestroy(h->side);
  if (h->ws) (void)hipFree(h->ws);
  for (auto& o : h->orders) (void)hipFree(o.d);
  if (h->d_flags) (void)hipFree(h->d_flags);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
  return MCPG_OK;
}

const char* mcpg_status_string(int status) {
  switch (status) {
    case MCPG_OK: return "ok";
    case MCPG_ERR_INVALID_ARG: return "invalid argument";
    case MCPG_ERR_SHAPE: return "shape mismatch";
    case MCPG_ERR_NONFINITE: return "non-finite input";
    case MCPG_ERR_OUT_OF_MEMORY: return "out of device memory";
    case MCPG_ERR_HIP: return "HIP runtime error";
    case MCPG_ERR_NO_DEVICE: return "no usable HIP device";
    case MCPG_ERR_UNSUPPORTED: return "unsupported size";
    default: return "unknown status";
  }
}

const char* mcpg_last_error(mcpg_handle h) { return h ? h->last_error.c_str() : "null handle"; }

int mcpg_synchronize(mcpg_handle h) {
  MCPG_CHECK_HANDLE(h);
  MCPG_HIP(h, hipSetDevice(h->device));
  MCPG_HIP(h, hipStreamSynchronize(h->stream));
  return MCPG_OK;
}

int mcpg_get_stream(mcpg_handle h, void** stream) {
  MCPG_CHECK_HANDLE(h);
  if (!stream) return MCPG_ERR_INVALID_ARG;
  *stream = reinterpret_cast<void*>(h->stream);
  return MCPG_OK;
}

int mcpg_last_launch_info(mcpg_handle h, int* launches, const char** dominant_kernel) {
  MCPG_CHECK_HANDLE(h);
  if (launches) *launches = h->last_launches;
  if (dominant_kernel) *dominant_kernel = h->last_kernel;
  return MCPG_OK;
}

int mcpg_set_profiling(mcpg_handle h, int enable) {
  MCPG_CHECK_HANDLE(h);
  h->profiling = enable != 0;
  return MCPG_OK;
}

int mcpg_last_timings(mcpg_handle h, int max_entries, const char** names, float* ms, int* count) {
  MCPG_CHECK_HANDLE(h);
  if (!count) return MCPG_ERR_INVALID_ARG;
  MCPG_HIP(h, hipSetDevice(h->device));
  MCPG_HIP(h, hipStreamSynchronize(h->stream));
  const int n = (int)h->timings.size();
  *count = n;
  for (int i = 0; i < n && i < max_entries; ++i) {
    if (names) names[i] = h->timings[i].name;
    if (ms) {
      float t = 0.0f;
      MCPG_HIP(h, hipEventElapsedTime(&t, h->timings[i].start, h->timings[i].stop));
      ms[i] = t;
    }
  }
  return MCPG_OK;
}

}  // extern "C"
