// Internal helpers shared by the HIP translation units of libmcpgpu.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <functional>
#include <string>
#include <thread>
#include <vector>

#include "mcpgpu/mcpgpu.h"

struct mcpg_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string last_error;
  void* ws = nullptr;        // grow-only device workspace
  size_t ws_bytes = 0;
  int* d_flags = nullptr;    // small device array: error flags (words 0-7), pair-queue counters (words 8-31)
  int last_launches = 0;
  const char* last_kernel = "";
  // per-launch timing of the most recent call (only when profiling is on)
  bool profiling = false;
  struct Timing {
    const char* name;
    hipEvent_t start, stop;
  };
  std::vector<Timing> timings;
  std::vector<hipEvent_t> event_pool;
  size_t events_used = 0;
  // cached device lists of upper-triangle tile coordinates (see mcpg::tile_order)
  struct TileOrder {
    uint32_t* d = nullptr;
    int64_t ni = -1, nj = -1, count = 0;
    int ratio = 0, gi = 0, gj = 0, xcd = 0;
  };
  std::vector<TileOrder> orders;
  // host copy of the tile list of the last block MI launch (kept alive for its upload)
  std::vector<uint32_t> block_order;
  // > 0 while a composite entry point (mcpg_mcpnet_*) runs its stages through other entry
  // points: the stages then keep the composite call's launch log and timings
  int nest = 0;
  // two pinned staging buffers for large host <-> device copies (see mcpg::copy_2d)
  void* pin[2] = {nullptr, nullptr};
  hipEvent_t pin_ev[2] = {nullptr, nullptr};
  bool pin_busy[2] = {false, false};
  // a second stream (default priority) with its own two pinned buffers, for a device -> host
  // download that overlaps the kernels queued after it on `stream` (see mcpg::Download)
  hipStream_t side = nullptr;
  void* side_pin[2] = {nullptr, nullptr};
  hipEvent_t side_ev[3] = {nullptr, nullptr, nullptr};   // the two buffers, and the start
  // max-min key kernel: key blocks of the split tail tiles, and resident workgroups per device
  int* mm_tail = nullptr;
  size_t mm_tail_bytes = 0;
  int mm_slots[4] = {0, 0, 0, 0};   // K chunk 16, 32; plain, then widened outputs
  int mm64_slots[2] = {0, 0}; // FP64 products, K chunk 8, 16 doubles
  // device address of the static scratch of the allocation-free entry point on this handle's
  // device (set by mcpg_create), and the error if it could not be taken
  void* scratch = nullptr;
  int scratch_err = 0;
};

namespace mcpg {

constexpr size_t kAlign = 256;
inline size_t align_up(size_t x, size_t a = kAlign) { return (x + a - 1) / a * a; }
__host__ __device__ inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

int set_error(mcpg_ctx* h, int status, const std::string& msg);
int hip_error(mcpg_ctx* h, hipError_t e, const char* where);

// Takes the device address of the allocation-free entry point's static scratch for the handle's
// device (current when called); on failure h->scratch stays null and h->scratch_err says why.
void init_scratch(mcpg_ctx* h);

// Ensures the workspace holds at least `bytes`; returns its base pointer (or nullptr).
void* workspace(mcpg_ctx* h, size_t bytes, int* status);

// Carves aligned sub-buffers out of a base pointer.
struct Carver {
  char* base;
  size_t off = 0;
  explicit Carver(void* b) : base(static_cast<char*>(b)) {}
  template <typename T>
  T* take(size_t count) {
    T* p = reinterpret_cast<T*>(base + off);
    off = align_up(off + count * sizeof(T));
    return p;
  }
};
// Size counter with the same alignment rule as Carver.
struct Sizer {
  size_t off = 0;
  template <typename T>
  void add(size_t count) { off = align_up(off + count * sizeof(T)); }
};

// 2-D strided copies between host/device buffers on the handle's stream.  Host <-> device copies
// of more than a few MB are staged through the handle's pinned buffers; a device -> host copy
// has completed when copy_2d returns.
// Pinned staging buffer b (0 or 1) of the handle for a host block of at most 32 MB (nullptr with
// *st == MCPG_OK when the block is larger), once the copy that last used it has finished;
// pinned_upload sends its first `bytes` to dst on the handle's stream.
void* pinned_staging(mcpg_ctx* h, int b, size_t bytes, int* st);
int pinned_upload(mcpg_ctx* h, int b, void* dst, size_t bytes);
// A device -> host copy of `rows` rows that runs on the handle's side stream, after the work
// queued on the handle's stream so far, through the side pinned buffers, driven by a host thread:
// start() returns at once, so later kernels on the handle's stream overlap it; wait() joins it
// and returns its status.  Copies of at most 8 MB are done synchronously by start().
struct Download {
  std::thread th;
  int status = MCPG_OK;
  int start(mcpg_ctx* h, void* dst, size_t dpitch, const void* src, size_t spitch, size_t width_bytes, size_t rows);
  int wait(mcpg_ctx* h);
  ~Download() {
    if (th.joinable()) th.join();
  }
};
int copy_2d(mcpg_ctx* h, void* dst, size_t dpitch, const void* src, size_t spitch, size_t width_bytes,
            size_t rows, hipMemcpyKind kind);

constexpr int kFlagWords = 32;   // ints in mcpg_ctx::d_flags

// Clears the device error flags and reads them back (synchronising the stream).
int reset_flags(mcpg_ctx* h);
int read_flags(mcpg_ctx* h, int* flags, int n);

int finish(mcpg_ctx* h, const char* where);  // stream sync + error check

// Device array of the tiles (ti, tj) of an ni x nj grid with tj >= ratio * ti, packed as
// (ti << 16) | tj and indexed by block id.  Tiles are ordered in gi x gj blocks (blocks row-major,
// tiles row-major inside a block), so that workgroups running together share gi row panels and
// gj column panels through L2 instead of one row panel and as many column panels as there are
// workgroups.  With xcd_deal set the order is further dealt so that each of the 8 XCDs
// (workgroups go to XCDs round-robin by id) runs one contiguous stretch of it.  Cached per handle.
const uint32_t* tile_order(mcpg_ctx* h, int64_t ni, int64_t nj, int ratio, int gi, int gj, int xcd_deal, int* status);

// The upper triangle (ti <= tj) of an nb x nb grid in group x group blocks.  Both kernels that use
// it measured faster without the XCD deal at 20000 genes.
const uint32_t* tri_order(mcpg_ctx* h, int64_t nb, int group, int xcd_deal, int* status);

// compute_maxmins on device matrices (M1 = A (x) A, M2 = A (x) M1, M3 = M1 (x) M1; M1 is
// required, M2 and M3 are formed when set).  `after(q, r0, r1)` runs once rows [r0, r1) of
// product q (0: M1, 1: M2, 2: M3) are on the handle's stream, before anything later is launched,
// so that their consumers (a ratio, a download on the side stream) overlap the next product or
// row panel: every product but the last whole, the last one (a triangle on the FP32-key kernel)
// in a few row panels.  A nonzero return ends the chain with that status.  It must not use the
// handle's workspace.
using ChainHook = std::function<int(int, int64_t, int64_t)>;
int maxmins_chain(mcpg_ctx* h, const float* A, int64_t N, float* M1, float* M2, float* M3, const ChainHook* after);
int maxmins_chain(mcpg_ctx* h, const double* A, int64_t N, double* M1, double* M2, double* M3, const ChainHook* after);

// Tile-group size: the value of environment variable `name` if set (1..256), else `def`.
int env_group(const char* name, int def);

// Starts an entry point: clears the launch counters and the timing list of the previous call.
void begin_call(mcpg_ctx* h);

// Brackets one kernel launch with HIP events on the handle's stream when profiling is on.
struct LaunchScope {
  mcpg_ctx* h;
  hipEvent_t stop = nullptr;
  LaunchScope(mcpg_ctx* hh, const char* name);
  ~LaunchScope();
};

// Maps a physical block id onto a logical tile id so that each XCD of the chip (blocks are
// dispatched round-robin over 8 XCDs) works on one contiguous range of logical tiles, which
// keeps the operand panels those tiles share inside that XCD's L2.
__device__ __forceinline__ int64_t xcd_remap(int64_t b, int64_t total) {
  constexpr int64_t kXcd = 8;
  const int64_t x = b % kXcd;
  const int64_t local = b / kXcd;
  const int64_t q = total / kXcd, r = total % kXcd;
  const int64_t prefix = x * q + (x < r ? x : r);
  return prefix + local;
}

}  // namespace mcpg

#define MCPG_CHECK_HANDLE(h)                                   \
  do {                                                         \
    if (!(h)) return MCPG_ERR_INVALID_ARG;                     \
  } while (0)

#define MCPG_HIP(h, call)                                      \
  do {                                                         \
    hipError_t e_ = (call);                                    \
    if (e_ != hipSuccess) return mcpg::hip_error((h), e_, #call); \
  } while (0)
